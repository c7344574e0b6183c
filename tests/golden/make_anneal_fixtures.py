"""Writes tests/golden/anneal_ref.json and tests/golden/metrics_ref.json.

anneal_ref.json: the reference's own CPU annealers -- LooperSolver::
MonteCarloHeatmap, MonteCarloArcs and MonteCarloArcsSmooth (LooperSolver.hpp
:1424-1565, :4670-4796, :4798-5047) compiled in place from the reference
sources into oracle/_ref/levels_ref by oracle/Makefile -- run on the seeded
plain workloads of pmmc_amd.synth.config_case over 16 seeds, each seeded with
srand(1 + seed) as the reference's `-j` option seeds glibc rand().  Recorded
per seed: the returned score, the number of moves, and the quality of the
result -- for the heatmap level the distance-map Pearson, SCC and RMSD to the
ground-truth polymer computed by the reference's own MetricsFramework /
fast_rmsd (the driver's `metrics` mode); for the arc level the Pearson
correlation of the anchor distances with their expected distances over the
arc pairs.

metrics_ref.json: the reference's metrics (MetricsFramework::pearsonCorrelation,
distanceCorrelation, structuralSimilarity, Chromosome::calcRMSD) on seeded
random structures and maps, for the metric restatements to be pinned against.

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_anneal_fixtures.py              # everything
    python tests/golden/make_anneal_fixtures.py --levels 3   # one level of
        anneal_ref.json, the other levels and metrics_ref.json kept
"""
from __future__ import annotations

import concurrent.futures as cf
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import metrics as OM  # noqa: E402
from oracle import refdriver as R  # noqa: E402
from pmmc_amd import synth  # noqa: E402

SEEDS = list(range(16))
# smooth instance of the band test: the 19,951-bead segment of the smooth
# workload, 400 fixed anchors with 49 beads between consecutive ones (about
# 39 M moves per chain at this schedule)
SMOOTH_ANCHORS, SMOOTH_ADD = 400, 49
HEADER = R.HEADER
hx = R.hx
run_ref = R.run


def ref_anneal(level: int, seed: int):
    prob, step, truth = synth.config_case(level, seed, SMOOTH_ANCHORS, SMOOTH_ADD)
    score, moves, pos = R.anneal(level, prob, step, synth.CONFIG_SETTINGS[level], 1 + seed)
    rec = {"seed": seed, "score": score, "moves": moves}
    if level == 1:
        rec.update(ref_metrics(pos, truth))
    if level == 2:
        rec["pearson_targets"] = arcs_target_pearson(pos, prob.exp_dist)
    return rec


def ref_metrics(recon, truth, h: int = 5, step: int = 1) -> dict:
    """The reference's own metrics of a reconstruction against the truth."""
    n = len(recon)
    text = "\n".join([f"metrics {HEADER}", f"{n} {h}", hx(recon), hx(truth),
                      hx(OM.distance_map(recon)), hx(OM.distance_map(truth)), str(step)]) + "\n"
    _, pear, dcor, scc, rmsd = run_ref(text, "@@metrics")[0].split()
    return {"pearson": float.fromhex(dcor), "scc": float.fromhex(scc),
            "rmsd": float.fromhex(rmsd), "map_pearson": float.fromhex(pear)}


def arcs_target_pearson(pos, exp_dist) -> float:
    """Pearson correlation of anchor distances with the expected distances of
    the arc pairs (i < j, e > 0)."""
    i, j = np.nonzero(np.triu(exp_dist > 1e-6, 1))
    d = np.linalg.norm(pos[i].astype(np.float64) - pos[j].astype(np.float64), axis=1)
    return OM.pearson(d.astype(np.float32), exp_dist[i, j])


def metrics_cases():
    cases = []
    for k, n in enumerate((2, 3, 5, 17, 40, 64, 130)):
        rng = np.random.default_rng(100 + k)
        a = rng.normal(size=(n, 3)).astype(np.float32) * np.float32(3.0)
        # b: a rotated, shifted, noisy copy (and one reflected case)
        q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        if k == 4:
            q = q @ np.diag([1.0, 1.0, -1.0])
        b = (a @ q.T + rng.normal(size=3) + 0.2 * rng.normal(size=(n, 3))).astype(np.float32)
        ha, hb = OM.distance_map(a), OM.distance_map(b)
        step = 1 + (k % 3)
        h = 5 if k % 2 == 0 else 2
        text = "\n".join([f"metrics {HEADER}", f"{n} {h}", hx(a), hx(b), hx(ha), hx(hb),
                          str(step)]) + "\n"
        _, pear, dcor, scc, rmsd = run_ref(text, "@@metrics")[0].split()
        cases.append({"n": n, "h": h, "step": step, "a": a.reshape(-1).tolist(),
                      "b": b.reshape(-1).tolist(),
                      "map_pearson": float.fromhex(pear), "distance_correlation": float.fromhex(dcor),
                      "scc": float.fromhex(scc), "rmsd": float.fromhex(rmsd)})
    return cases


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--levels", type=int, nargs="*", default=[1, 2, 3])
    levels = ap.parse_args(argv).levels
    if not R.available():
        print(f"{R.REF} missing: run `make -C oracle ref` (needs the reference sources)")
        return 1
    path = os.path.join(HERE, "anneal_ref.json")
    if sorted(levels) == [1, 2, 3]:
        out = {"metrics": metrics_cases()}
        json.dump(out, open(os.path.join(HERE, "metrics_ref.json"), "w"), indent=0)
        old = {}
    else:
        old = json.load(open(path))
    res = {"generator": "tests/golden/make_anneal_fixtures.py",
           "annealers": "LooperSolver::MonteCarloHeatmap / MonteCarloArcs / MonteCarloArcsSmooth "
                        "(reference, compiled in place), glibc rand() seeded with srand(1 + seed)",
           "settings": {str(k): v for k, v in synth.CONFIG_SETTINGS.items()},
           "smooth_anchors": SMOOTH_ANCHORS, "smooth_add": SMOOTH_ADD}
    jobs = [(lvl, s) for lvl in levels for s in SEEDS]
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        recs = list(ex.map(lambda t: (t[0], ref_anneal(*t)), jobs))
    for lvl in (1, 2, 3):
        res[str(lvl)] = [r for lv, r in recs if lv == lvl] if lvl in levels else old[str(lvl)]
    json.dump(res, open(os.path.join(HERE, "anneal_ref.json"), "w"), indent=1)
    for lvl in (1, 2, 3):
        sc = [r["score"] for r in res[str(lvl)]]
        mv = [r["moves"] for r in res[str(lvl)]]
        print(f"level {lvl}: score mean {np.mean(sc):.6g} sd {np.std(sc):.3g}, "
              f"moves mean {np.mean(mv):.0f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
