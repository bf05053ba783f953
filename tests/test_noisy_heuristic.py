"""Pins the oracle's noisy heuristic (OR_POLICY_HEURISTIC_NOISY, oracle/oracle.c) on the CPU.

The reference's heuristic drivers add an N(0, 0.11) draw to each move's weight and play the
largest sum (src/heuristicPlayout.cpp:16,26-36; src/heuristicPlayout.cu:58; src/heuristic.hpp:5).
Their random sources (a default-seeded std::default_random_engine shared by OpenMP threads,
cuRAND) cannot be replayed, so the restatement draws its noise from the ply's Philox word by a
Box-Muller transform (a ply whose moves all weigh the same plays r mod n instead: the argmax of
iid draws is uniform there), and parity is pinned in three layers:
  * the draws: the single-precision transform against the float64 closed form, and its
    distribution against N(0, 1) (moments, Kolmogorov-Smirnov, independence of a pair);
  * the mechanics (move order, weights, strict > from -infinity, the mover-order draw index):
    playout traces written by the reference's own move generation and weights with the same
    draws (tests/golden/noisy.npz, traces);
  * the distribution: per-ply choice counts of the reference's weights with its own
    std::normal_distribution noise, and winner counts of the reference's own
    HostHeuristicPlayoutDriver::runPlayouts, against the oracle's, by chi-square tests.  The
    noise-free heuristic and the oracle rebuilt with sigma 0.22 (OR_MUTATION 7) must fail them.
Outcome counts move little with the noise (tied weights, the common case, give a uniform choice
either way); the per-ply counts at positions whose weights differ are what tell them apart.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest
from scipy import stats

from oracle import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NOISY = pyoracle.POLICY_HEURISTIC_NOISY
P_MIN = 1e-3  # a fixed-seed sample: a pass is deterministic; a correct oracle gives p >> P_MIN


def fmix(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x85EBCA6B) & 0xFFFFFFFF
    x ^= x >> 13
    x = (x * 0xC2B2AE35) & 0xFFFFFFFF
    x ^= x >> 16
    return x.astype(np.uint32)


def box_muller_f64(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """The transform's definition in double precision (numpy's log, cos, sin)."""
    u = ((a >> 8).astype(np.float64) + 1.0) / 2.0 ** 24
    ang = (b >> 30).astype(np.float64) * np.pi / 2 + \
        (((b >> 6) & 0xFFFFFF).astype(np.float64) / 2.0 ** 24 - 0.5) * np.pi / 2
    r = np.sqrt(-2.0 * np.log(u))
    return np.stack([r * np.cos(ang), r * np.sin(ang)], axis=1)


def pooled_chi2(pairs) -> tuple[float, int, float]:
    """Sum of two-sample chi-square statistics over positions (cells empty in both dropped)."""
    chi, dof = 0.0, 0
    for a, b in pairs:
        t = np.stack([a, b]).astype(np.float64)
        t = t[:, t.sum(axis=0) > 0]
        if t.shape[1] < 2:
            assert np.array_equal(t[0], t[1]), "a forced choice differs"
            continue
        r = stats.chi2_contingency(t, correction=False)
        chi += r.statistic
        dof += r.dof
    return chi, dof, float(stats.chi2.sf(chi, dof))


@pytest.fixture(scope="module")
def sigma_mutant(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("noisy") / "liboracle_m7.so")
    subprocess.run(["gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared", "-DOR_MUTATION=7",
                    os.path.join(ROOT, "oracle", "oracle.c"), "-o", out, "-lm"], check=True)
    return pyoracle.Oracle(out)


# ---------------------------------------------------------------------------------------------
# the draws
# ---------------------------------------------------------------------------------------------

def test_box_muller_matches_closed_form(oracle):
    rng = np.random.default_rng(11)
    a = rng.integers(0, 2 ** 32, 400_000, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2 ** 32, 400_000, dtype=np.uint64).astype(np.uint32)
    # the extremes of u and of the angle
    a[:4] = [0, 0xFFFFFFFF, 0x000000FF, 0xFFFFFF00]
    b[:4] = [0, 0xFFFFFFFF, 0x40000000, 0xBFFFFFC0]
    z = oracle.normal_pairs(a, b).astype(np.float64)
    assert np.abs(z - box_muller_f64(a, b)).max() < 2e-6
    assert z[1, 0] == 0.0 and z[1, 1] == 0.0  # u = 1: radius 0


def test_draws_are_standard_normal(oracle):
    rng = np.random.default_rng(12)
    a = rng.integers(0, 2 ** 32, 1_000_000, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2 ** 32, 1_000_000, dtype=np.uint64).astype(np.uint32)
    z = oracle.normal_pairs(a, b).astype(np.float64)
    for k in range(2):
        x = z[:, k]
        assert abs(x.mean()) < 5e-3 and abs(x.std() - 1.0) < 5e-3
        assert abs(stats.skew(x)) < 0.02 and abs(stats.kurtosis(x)) < 0.03
        assert stats.kstest(x, "norm").pvalue > P_MIN
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 5e-3
    assert abs(np.corrcoef(z[:, 0] ** 2, z[:, 1] ** 2)[0, 1]) < 5e-3


def test_move_draws_scale_and_pairing(oracle):
    """A ply's draws: sigma 0.11, moves 2j and 2j+1 share the pair of words j, and draws of
    different moves and plies are uncorrelated."""
    rng = np.random.default_rng(13)
    rs = rng.integers(0, 2 ** 32, 20000, dtype=np.uint64).astype(np.uint32)
    d = np.array([[oracle.heuristic_noise(int(r), m) for m in range(6)] for r in rs])
    assert abs(d.std() - 0.11) < 0.002
    c = np.corrcoef(d.T)
    assert np.abs(c - np.eye(6)).max() < 0.03
    r = rs[:64]
    for j in range(3):
        a = fmix(r + np.uint32((2 * j + 1) * 0x9E3779B9 & 0xFFFFFFFF))
        b = fmix(r + np.uint32((2 * j + 2) * 0x9E3779B9 & 0xFFFFFFFF))
        z = oracle.normal_pairs(a, b)
        for k in range(2):
            got = np.array([oracle.heuristic_noise(int(x), 2 * j + k) for x in r], dtype=np.float32)
            assert np.array_equal(got, z[:, k] * np.float32(0.11))


# ---------------------------------------------------------------------------------------------
# the mechanics: traces over the reference's move generation and weights
# ---------------------------------------------------------------------------------------------

def test_noisy_traces_match_reference_golden(oracle, golden):
    g = golden["noisy"]
    _, tr = oracle.run_playouts(g["trace_pos"], per_pos=16, seed=12345, epoch=0, policy=NOISY)
    exp = g["traces"]
    for f in ("winner", "plies", "hash", "max_n"):
        assert np.array_equal(tr[f], exp[f]), f


def test_noisy_differs_from_argmax(oracle, golden):
    """The two heuristics play the same index r mod n where every weight is equal, so a game
    changes only at a ply whose weights differ: a large share of the traces must differ."""
    g = golden["noisy"]
    _, tr = oracle.run_playouts(g["trace_pos"], per_pos=16, seed=12345, epoch=0,
                                policy=pyoracle.POLICY_HEURISTIC)
    assert np.mean(tr["hash"] != g["traces"]["hash"]) > 0.25


@pytest.mark.skipif(not pyoracle.reference_available(pyoracle.REF_SO),
                    reason="oracle/_ref not built here")
def test_noisy_traces_match_reference_build(oracle):
    ref = pyoracle.Reference(pyoracle.REF_SO)
    pos = oracle.start_positions(31, 128, 0, 70, True)
    _, a = oracle.run_playouts(pos, per_pos=32, seed=5, epoch=2, policy=NOISY)
    _, b = ref.run_playouts(pos, per_pos=32, seed=5, epoch=2, policy=NOISY)
    for f in ("winner", "plies", "hash", "max_n"):
        assert np.array_equal(a[f], b[f]), f


# ---------------------------------------------------------------------------------------------
# the distribution: against the reference's own noise and its own driver
# ---------------------------------------------------------------------------------------------

def choice_pairs(lib, g, policy):
    trials = int(g["choice_trials"])
    for i, p in enumerate(g["choice_pos"]):
        n = int(g["choice_n"][i])
        yield g["choice_counts"][i, :n], lib.choice_counts(p, policy, 4000 + i, trials)


def test_choice_frequencies_match_reference_noise(oracle, golden):
    g = golden["noisy"]
    chi, dof, p = pooled_chi2(choice_pairs(oracle, g, NOISY))
    assert dof > 50 and p > P_MIN, (chi, dof, p)


def test_equal_weights_choose_uniformly(oracle, golden):
    """Where every move weighs the same the reference's choice is the argmax of its draws alone;
    its counts and the oracle's r mod n counts must both be uniform."""
    g = golden["noisy"]
    lo = int(g["choice_single_from"])
    assert len(g["choice_pos"]) - lo >= 8
    trials = int(g["choice_trials"])
    chi_ref = chi_mine = 0.0
    dof = 0
    for i in range(lo, len(g["choice_pos"])):
        n = int(g["choice_n"][i])
        ref = g["choice_counts"][i, :n].astype(np.float64)
        mine = oracle.choice_counts(g["choice_pos"][i], NOISY, 77 + i, trials).astype(np.float64)
        chi_ref += float(((ref - trials / n) ** 2 / (trials / n)).sum())
        chi_mine += float(((mine - trials / n) ** 2 / (trials / n)).sum())
        dof += n - 1
    assert stats.chi2.sf(chi_ref, dof) > P_MIN and stats.chi2.sf(chi_mine, dof) > P_MIN


def test_choice_frequencies_reject_wrong_noise(oracle, golden, sigma_mutant):
    g = golden["noisy"]
    for lib, policy in ((oracle, pyoracle.POLICY_HEURISTIC), (sigma_mutant, NOISY)):
        _, _, p = pooled_chi2(choice_pairs(lib, g, policy))
        assert p < 1e-12, policy


def test_infinite_weights_take_the_first_in_reference_order(oracle, golden):
    """Moves that take the last opponent piece weigh +inf; inf plus a draw stays inf, so the
    strict > keeps the first of them in reference order, for either side to move."""
    g = golden["noisy"]
    lo, hi = int(g["choice_infinite_from"]), int(g["choice_single_from"])
    assert hi - lo >= 8
    turns = set()
    for i in range(lo, hi):
        p = g["choice_pos"][i]
        n = int(g["choice_n"][i])
        ref_counts = g["choice_counts"][i, :n]
        assert np.count_nonzero(ref_counts) == 1
        mine = oracle.choice_counts(p, NOISY, 9, 500)
        assert np.argmax(mine) == np.argmax(ref_counts) and mine.max() == 500
        turns.add(int(p["turn"]))
    assert turns == {0, 1}


def driver_pairs(lib, g, policy, per):
    w, _ = lib.run_playouts(g["driver_pos"], per_pos=per, seed=321, epoch=0, policy=policy,
                            traces=False)
    w = w.reshape(len(g["driver_pos"]), per)
    mine = np.stack([(w == 0).sum(1), (w == 1).sum(1), (w == -1).sum(1)], axis=1)
    return zip(g["driver_counts"], mine)


def test_outcomes_match_reference_driver(oracle, golden):
    """Winner counts against the reference's own HostHeuristicPlayoutDriver (16 midgame
    positions x 16384 playouts each, its default-seeded noise)."""
    g = golden["noisy"]
    chi, dof, p = pooled_chi2(driver_pairs(oracle, g, NOISY, 16384))
    assert dof >= 24 and p > P_MIN, (chi, dof, p)
    _, _, p_random = pooled_chi2(driver_pairs(oracle, g, pyoracle.POLICY_RANDOM, 16384))
    assert p_random < 1e-12
