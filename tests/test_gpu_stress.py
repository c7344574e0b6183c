"""Stress refinement on the GPU (pmmc_stress_refine, pmmc_tree_stress_refine
<- LooperSolver::stressRefineAnchors, LooperSolver.hpp:5638-5915).

* the device arc cosine equals the C library's acosf (the reference's) on
  every float in [-1, 1] (checksums over 2^31 inputs, libm on the host);
* pmmc_tree_stress_refine on every case of tests/golden/stress_ref.json
  leaves the positions the reference's own function leaves, bit for bit;
* pmmc_stress_refine equals the oracle's descent bit for bit on random
  problems of 2-3,000 beads, with the tree-summed decisions and with the
  in-order sums alike, one problem per call and many per launch; the unstable
  step case stops at the step floor as the oracle does.
"""
import json
import os
import time

import numpy as np
import pytest

from oracle import binding as O
from pmmc_amd import api
from pmmc_amd import tree as T

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FX = json.load(open(os.path.join(HERE, "golden", "stress_ref.json")))
STOP = {"max_iter": 0, "converged": 1, "lr": 2}
RENAME = {"stretching": "heat_stretching"}


def unhex(v):
    return np.array([float.fromhex(x) for x in v], np.float32)


def test_device_acosf_equals_libm_on_every_float(ctx):
    chunk = 1 << 27
    ranges = [(0, 0x3f800000 + 1), (0x80000000, 0x80000000 + 0x3f800000 + 1)]
    done = 0
    for lo, hi in ranges:
        for first in range(lo, hi, chunk):
            count = min(chunk, hi - first)
            dev = ctx.probe_acosf(first, count)
            host = O.acosf_sum(first, count, threads=16)
            assert dev == host, hex(first)
            done += count
    print(f"acosf: {done} inputs checked", flush=True)
    assert done == 2 * (0x3f800000 + 1)


@pytest.mark.parametrize("k", range(len(FX["cases"])))
def test_tree_stress_matches_reference(ctx, k):
    c = FX["cases"][k]
    p = FX["params"][c["params"]]
    chrs = [dict(name="chr1", anchors=c["anchors"], arcs=c["arcs"], long_arcs=c["long_arcs"])]
    pt = T.Tree(chrs, c["splits"])
    first, _ = pt.chromosome(0)
    n = len(c["anchors"])
    ids = list(range(first, first + n))
    pt.set_positions(ids, unhex(c["pos"]).reshape(n, 3))
    lp = T.level_params(**{RENAME.get(k2, k2): v for k2, v in p.items()})
    st = T.stress_refine(ctx, pt, 0, lp, c["k_stretch"], c["k_squeeze"], c["k_angular"],
                         max_iter=c["iters"], lr=c["lr"])
    got = pt.positions(ids)
    want = unhex(c["out"]).reshape(n, 3)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), c["seed"]
    assert st[0]["stop"] == STOP[c["stop"]]


def _problem(rng, n, pairs_per_bead=3, ka=0.1):
    pos = np.cumsum(rng.normal(0, 2.0, size=(n, 3)), 0).astype(np.float32)
    bb = rng.uniform(1.5, 4.0, size=max(n - 1, 1)).astype(np.float32)
    bb[rng.random(bb.size) < 0.05] = -1.0          # some springs off
    pairs, seen = [], set()
    for _ in range(pairs_per_bead * n):
        i = int(rng.integers(0, n))
        j = int(min(n - 1, i + rng.integers(2, 12)))
        if i == j or (i, j) in seen:
            continue
        seen.add((i, j))
        pairs.append((i, j, float(np.float32(rng.uniform(1.0, 6.0)))))
    return dict(pos=pos, backbone=bb, pairs=pairs, k_stretch=0.1, k_squeeze=0.1, k_angular=ka)


def _oracle(q, iters, lr):
    return O.stress_refine(q["pos"], q["backbone"], q["pairs"], q["k_stretch"], q["k_squeeze"],
                           q["k_angular"], iters, lr)


@pytest.mark.parametrize("n,iters,lr", [(2, 50, 0.05), (3, 80, 0.05), (57, 400, 0.05),
                                        (700, 600, 0.02), (3000, 300, 0.01)])
def test_flat_stress_matches_oracle(ctx, n, iters, lr):
    rng = np.random.default_rng(n)
    q = _problem(rng, n)
    print(f"stress n={n} pairs={len(q['pairs'])} iters={iters}", flush=True)
    t0 = time.perf_counter()
    want, wst = _oracle(q, iters, lr)
    t1 = time.perf_counter()
    (got, st), = ctx.stress_refine([q], iters, lr)
    t2 = time.perf_counter()
    (got2, st2), = ctx.stress_refine([q], iters, lr, in_order_sums=True)
    print(f"oracle {1e3 * (t1 - t0):.1f} ms, device {1e3 * (t2 - t1):.1f} ms, "
          f"{st['iterations']} updates, stop {st['stop']}, in-order sums needed "
          f"{st['exact_sums']}", flush=True)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(got2.view(np.uint32), want.view(np.uint32))
    for key in ("iterations", "stop"):
        assert st[key] == wst[key] == st2[key], key
    assert np.float32(st["lr"]) == np.float32(wst["lr"])
    for key in ("total", "backbone", "pairs", "angular"):   # the in-order sums, as printed
        assert st[key] == wst[key] == st2[key], key
    assert st2["exact_sums"] >= st["iterations"]
    if n > 2:
        assert np.abs(got - q["pos"]).max() > 0.1


def test_many_problems_one_launch(ctx):
    rng = np.random.default_rng(99)
    qs = [_problem(rng, n, ka=ka) for n, ka in ((40, 0.1), (5, 0.0), (300, 0.3), (1, 0.1),
                                                 (120, 0.1), (2, 0.1))]
    res = ctx.stress_refine(qs, 500, 0.05)
    for q, (got, st) in zip(qs, res):
        want, wst = _oracle(q, 500, 0.05)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), len(q["pos"])
        assert st["stop"] == wst["stop"] and st["iterations"] == wst["iterations"]
    # a single bead has nothing to do
    assert res[3][1]["stop"] == 3 and np.array_equal(res[3][0], qs[3]["pos"])


def test_unstable_step_stops_at_the_floor(ctx):
    q = dict(pos=np.array([[0, 0, 0], [1, 0, 0]], np.float32), backbone=np.array([-1.0], np.float32),
             pairs=[(0, 1, 1e-3)], k_stretch=0.0, k_squeeze=0.0, k_angular=0.0)
    (got, st), = ctx.stress_refine([q], 100, 1.5e-5)
    want, wst = _oracle(q, 100, 1.5e-5)
    assert st["stop"] == wst["stop"] == 2 and st["iterations"] == wst["iterations"] == 4
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_bad_pairs_rejected(ctx):
    q = dict(pos=np.zeros((4, 3), np.float32), backbone=np.ones(3, np.float32),
             pairs=[(0, 4, 1.0)])
    with pytest.raises(api.PmmcError):
        ctx.stress_refine([q], 10, 0.05)
    q["pairs"] = [(2, 2, 1.0)]
    with pytest.raises(api.PmmcError):
        ctx.stress_refine([q], 10, 0.05)
