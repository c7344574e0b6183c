"""Writes tests/golden/stress_ref.json: seeded inputs of
LooperSolver::stressRefineAnchors (LooperSolver.hpp:5638-5915) and the
positions the reference's own function leaves.

The outputs come from oracle/_ref/levels_ref, mode "stress", which
oracle/Makefile compiles from the reference sources in place
(oracle/ref_levels_driver.cpp): a tree of one chromosome (anchors, arcs with
chromosome-local indices, segment splits), long arcs in genomic positions,
spring constants, iterations and step, and the anchors' start positions.
Cases cover: sizes 12-400 anchors, the step halving (a large step), the
backbone off (adjacent pairs kept), bending off, duplicated pairs between
arcs and long arcs, the count-to-distance power model, and anchors given out
of genomic order relative to their arcs.  Floats are stored as float32 bit
patterns (hex), so the fixture is exact.  Run where the reference sources
exist:
    make -C oracle ref && python tests/golden/make_stress_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

PARAMS = dict(count_a=0.5, count_scale=20.0, count_shift=1.0, count_base=0.01,
              heat_scale=100.0, heat_power=-0.333, heat_scale_inter=100.0,
              heat_power_inter=-1.0, stretching=2.0, glen_base=0.0, glen_scale=1.0,
              glen_power=0.5)
PARAMS_POWER = dict(PARAMS, count_a=-0.333, count_scale=2.5, count_shift=-2.0, count_base=0.05,
                    glen_base=0.1, glen_scale=0.02, glen_power=0.6)
# pair targets below 1e-6: no pairs at all (only springs and bending)
PARAMS_NOPAIRS = dict(PARAMS, count_base=0.0, count_scale=1e-9)

# (seed, anchors, iterations, lr, k_stretch, k_squeeze, k_angular, long arcs, params)
CASES = [
    (1, 12, 50, 0.05, 0.1, 0.1, 0.1, 4, "default"),
    (2, 40, 300, 0.05, 0.1, 0.1, 0.1, 10, "default"),
    (3, 80, 1000, 0.05, 0.1, 0.1, 0.1, 20, "default"),
    (4, 150, 2000, 0.02, 0.1, 0.1, 0.1, 40, "default"),
    (5, 60, 500, 5.0, 0.1, 0.1, 0.1, 10, "default"),
    (6, 60, 500, 0.05, 0.0, 0.0, 0.1, 10, "default"),
    (7, 60, 500, 0.05, 0.1, 0.1, 0.0, 10, "default"),
    (8, 400, 3000, 0.01, 1.0, 0.5, 0.1, 120, "default"),
    (9, 90, 800, 0.05, 0.1, 0.1, 0.1, 30, "power"),
    (10, 30, 400, 0.05, 0.0, 0.0, 1.0, 0, "nopairs"),   # straight start: the energy rises
]


def f32(x):
    return float(np.float32(x))


def hx(a):
    return [float(v).hex() for v in np.asarray(a, np.float32).reshape(-1)]


def make_case(seed, n, n_long):
    """One chromosome: anchors 4-30 kb apart, arcs spanning up to 6 anchors
    with counts 2-6 (targets of a few units at the default count model),
    long arcs between random anchors (some repeating an arc's pair), and a
    random-walk start."""
    rng = np.random.default_rng(seed)
    gaps = rng.integers(4000, 30000, size=n)
    lens = rng.integers(500, 3000, size=n)
    starts = 10_000 + np.cumsum(gaps + lens) - lens
    anchors = [(int(s), int(s + l)) for s, l in zip(starts, lens)]
    pairs = set()
    for _ in range(n):
        a = int(rng.integers(0, n - 1))
        b = int(min(n - 1, a + rng.integers(1, 7)))
        pairs.add((a, b))
    arcs = [(a, b, int(rng.integers(2, 7))) for a, b in sorted(pairs)]
    longs = []
    for k in range(n_long):
        if k % 4 == 0 and arcs:            # the pair of an existing arc: the arc wins
            a, b, _ = arcs[int(rng.integers(0, len(arcs)))]
        else:
            a, b = sorted(int(x) for x in rng.integers(0, n, size=2))
        x = int(rng.integers(anchors[a][0], anchors[a][1] + 1))
        y = int(rng.integers(anchors[b][0], anchors[b][1] + 1))
        longs.append((x, y, int(rng.integers(2, 7))))
    lo, hi = anchors[0][0] - 1000, anchors[-1][1] + 1000
    cut = int(rng.integers(lo + 1, hi))
    splits = [("chr1", lo, cut), ("chr1", cut, hi)]
    pos = np.cumsum(rng.normal(0, 2.0, size=(n, 3)), 0).astype(np.float32)
    return dict(anchors=anchors, arcs=arcs, long_arcs=longs, splits=splits), pos


def run(case, pos, params, ks, kq, ka, iters, lr):
    from oracle import refdriver
    _, order, out, log = refdriver.stress(case, params, ks, kq, ka, iters, lr, pos)
    return order, out, log


def main():
    fx = dict(params=dict(default=PARAMS, power=PARAMS_POWER, nopairs=PARAMS_NOPAIRS), cases=[])
    for seed, n, iters, lr, ks, kq, ka, n_long, pk in CASES:
        case, pos = make_case(seed, n, n_long)
        if pk == "nopairs":    # collinear, evenly spaced along a fixed direction
            u = np.array([0.48, -0.6, 0.64], np.float32)
            pos = (np.arange(n, dtype=np.float32)[:, None] * np.float32(3.0) * u).astype(np.float32)
        order, out, log = run(case, pos, fx["params"][pk], ks, kq, ka, iters, lr)
        assert np.isfinite(out).all(), seed
        stop = ("converged" if any("converged" in ln for ln in log) else
                "lr" if any("stopping" in ln for ln in log) else "max_iter")
        fx["cases"].append(dict(seed=seed, params=pk, k_stretch=ks, k_squeeze=kq, k_angular=ka,
                                iters=iters, lr=lr, anchors=case["anchors"], arcs=case["arcs"],
                                long_arcs=case["long_arcs"], splits=case["splits"],
                                pos=hx(pos), order=order, out=hx(out), stop=stop))
        print(seed, n, "stop", stop, "moved", float(np.abs(out - pos).max()), flush=True)
    with open(os.path.join(HERE, "stress_ref.json"), "w") as f:
        json.dump(fx, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
