"""Writes tests/golden/singleton_ref.json: seeded inputs for the heatmap input
construction upstream of createDistanceHeatmap, and what the reference's own
host code returns for them (oracle/_ref/levels_ref, compiled in place from
the reference sources by oracle/Makefile; oracle/ref_levels_driver.cpp):
  singleton  LooperSolver::createSingletonHeatmap (LooperSolver.hpp:2517-2749)
             on the clusters of one level of a tree built by createTreeGenome,
             with singleton records, long-range arcs, a selected region and
             centromeres
  normalize  normalizeHeatmap (:2996-3040), normalizeHeatmapDiagonalTotal
             (:3146-3172), normalizeHeatmapInter (:3106-3144)
Floats are stored as float32 bit patterns (base64), so the fixture is exact.
Run where the reference sources exist:
    make -C oracle ref && python tests/golden/make_singleton_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_levels_fixtures as L  # noqa: E402  (tree inputs, driver I/O helpers)


def tree_out(t):
    return L.parse_tree(L.run(L.param_line("tree", L.PARAMS_DEFAULT) + L.tree_text(t)))


def level_clusters(t, tout, level):
    """[(chr, [(start, end) of the level's clusters in level order])]"""
    out = []
    for c in t["chrs"]:
        ids = tout["levels"][c["name"]][str(level)]
        out.append((c["name"], [(tout["clusters"][k]["start"], tout["clusters"][k]["end"]) for k in ids]))
    return out


def records(rng, lvl, n, extra_chr="chrM"):
    """Singleton records (chr1 sta stb chr2 enda endb score) over the level's
    clusters: intra and inter pairs, a few on an unknown chromosome, negative
    starts, starts exactly on bin breaks, and same-bin pairs."""
    names = [c for c, _ in lvl]
    span = {c: (cl[0][0] - 2000, cl[-1][1] + 2000) for c, cl in lvl}
    breaks = {}
    for c, cl in lvl:
        breaks[c] = [(cl[i][1] + cl[i + 1][0]) // 2 for i in range(len(cl) - 1)]
    out = []
    for k in range(n):
        c1 = names[int(rng.integers(len(names)))]
        c2 = c1 if rng.random() < 0.7 else names[int(rng.integers(len(names)))]
        lo1, hi1 = span[c1]
        lo2, hi2 = span[c2]
        sta = int(rng.integers(lo1, hi1))
        enda = int(rng.integers(lo2, hi2))
        r = rng.random()
        if r < 0.04 and breaks[c1]:
            sta = int(breaks[c1][int(rng.integers(len(breaks[c1])))])   # exactly on a break
        elif r < 0.06:
            sta = -int(rng.integers(1, 500))                            # negative start
        elif r < 0.09:
            enda = sta + int(rng.integers(0, 50))                       # same bin (mostly)
            c2 = c1
        stb = sta + int(rng.integers(20, 300))
        endb = enda + int(rng.integers(20, 300))
        if rng.random() < 0.02:
            c1 = extra_chr                                              # not processed
        out.append((c1, sta, stb, c2, enda, endb, int(rng.integers(1, 20))))
    return out


def long_arcs(rng, lvl, n):
    out = []
    for c, cl in lvl:
        lo, hi = cl[0][0], cl[-1][1]
        for _ in range(n):
            a = int(rng.integers(lo, hi))
            b = int(rng.integers(lo, hi))
            out.append((c, min(a, b), max(a, b), int(rng.integers(2, 40))))
    return out


def run_singleton(t, level, diag, scale, power, region, cents, longs, recs):
    text = (L.param_line("singleton", L.PARAMS_DEFAULT) + L.tree_text(t) +
            f"{level} {diag} {L.fstr(scale)} {L.fstr(power)}\n{region[0]} {region[1]}\n" +
            f"{len(cents)}\n" + "".join(f"{c} {s} {e}\n" for c, s, e in cents) +
            f"{len(longs)}\n" + "".join(f"{c} {s} {e} {sc}\n" for c, s, e, sc in longs) +
            f"{len(recs)}\n" + "".join(" ".join(str(v) for v in r) + "\n" for r in recs))
    lines = L.run(text).splitlines()
    k = next(i for i, l in enumerate(lines) if l.startswith("@@singleton "))
    n = int(lines[k].split()[1])
    return np.array([[L.hexf(x) for x in l.split()] for l in lines[k + 1:k + 1 + n]], np.float32)


def main():
    if not os.path.exists(L.BIN):
        raise SystemExit("oracle/_ref/levels_ref not built (make -C oracle ref; needs the reference sources)")
    rng = np.random.default_rng(20261205)
    trees = L.tree_cases(np.random.default_rng(20261204))  # the tree inputs of levels_ref.json
    fx = {"source": "reference LooperSolver host code, oracle/_ref/levels_ref", "singleton": [],
          "normalize": []}
    cases = [
        # (tree, level, diag, region, centromere spec, n records, long arcs per chr, scale, power)
        (2, 1, 1, (0, 0), "none", 400, 3, 10.0, 2.0),
        (2, 2, 1, (0, 0), "mid", 1500, 4, 10.0, 2.0),
        (4, 1, 2, (0, 0), "edges", 800, 2, 3.0, 0.5),
        (1, 3, 1, "inner", "none", 2500, 5, 10.0, 2.0),
        (4, 3, 2, (0, 0), "mid", 4000, 6, 1.5, 1.0),
    ]
    for ti, level, diag, region, cspec, nrec, nlong, scale, power in cases:
        t = trees[ti]
        tout = tree_out(t)
        lvl = level_clusters(t, tout, level)
        if region == "inner":
            lo, hi = lvl[0][1][0][0], lvl[0][1][-1][1]
            region = (lo + (hi - lo) // 5, hi - (hi - lo) // 5)
        cents = []
        for c, cl in lvl:
            lo, hi = cl[0][0], cl[-1][1]
            if cspec == "mid":      # inside the chromosome: bins before, across and after it
                a = lo + (hi - lo) // 3
                cents.append((c, a, a + (hi - lo) // 6))
            elif cspec == "edges":  # covering the first bins / the last bins
                cents.append((c, lo - 5000, lo + (hi - lo) // 4) if c == lvl[0][0]
                             else (c, hi - (hi - lo) // 4, hi + 5000))
        longs = long_arcs(rng, lvl, nlong)
        recs = records(rng, lvl, nrec)
        heat = run_singleton(t, level, diag, scale, power, region, cents, longs, recs)
        fx["singleton"].append(dict(
            tree=ti, level=level, diag=diag, region=list(region), centromeres=cents,
            long_arcs=longs, long_scale=scale, long_power=power,
            chrs=[c for c, _ in lvl], clusters=[[list(x) for x in cl] for _, cl in lvl],
            records=recs, n=int(heat.shape[0]), out=L.b64(heat)))
    for sizes, val, scale in [([6], 1.0, 2.0), ([5, 9], 1.0, 2.0), ([4, 7, 3], 0.5, 1.5),
                              ([12, 20], 2.0, 0.75)]:
        n = sum(sizes)
        h = rng.uniform(0, 50, size=(n, n)).astype(np.float32)
        h = ((h + h.T) / 2).astype(np.float32)
        np.fill_diagonal(h, 0.0)
        if n > 20:
            h[3, :] = 0.0   # an empty row (its normalisation factor is 0)
            h[:, 3] = 0.0
        text = (L.param_line("normalize", L.PARAMS_DEFAULT) + f"{len(sizes)} " +
                " ".join(str(k) for k in sizes) + f"\n{L.fstr(val)} {L.fstr(scale)}\n" +
                "".join(" ".join(L.fstr(x) for x in row) + "\n" for row in h))
        lines = L.run(text).splitlines()
        k = next(i for i, l in enumerate(lines) if l.startswith("@@normalize "))
        mats = [np.array([[L.hexf(x) for x in l.split()] for l in lines[k + 1 + m * n:k + 1 + (m + 1) * n]],
                         np.float32) for m in range(3)]
        fx["normalize"].append(dict(sizes=sizes, val=val, scale=scale, heat=L.b64(h),
                                    normalized=L.b64(mats[0]), diagonal_total=L.b64(mats[1]),
                                    inter=L.b64(mats[2])))
    path = os.path.join(HERE, "singleton_ref.json")
    with open(path, "w") as fh:
        json.dump(fx, fh, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    sys.exit(main())
