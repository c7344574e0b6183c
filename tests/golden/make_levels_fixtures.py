"""Writes tests/golden/levels_ref.json: seeded inputs for the host-side rows of
the annealing path and what the reference's own host code returns for them.

The outputs come from oracle/_ref/levels_ref, which oracle/Makefile compiles
from the reference sources in place (oracle/ref_levels_driver.cpp):
  tree      LooperSolver::createTreeGenome / findGaps / findSplit / setLevel
            (LooperSolver.hpp:2119-2452, :3087-3104)
  densify   LooperSolver::densifyActiveRegion (:3836-3898) and the
            dist_to_next / noise_size rule of the arc driver (:4339-4354)
  expdist   LooperSolver::calcAnchorExpectedDistancesHeatmap (:5494-5627)
  distheat  LooperSolver::createDistanceHeatmap (:3042-3085)
  score     calcScoreHeatmapActiveRegion / calcScoreDistancesActiveRegion /
            calcScoreStructureSmooth (:3583-3612, :3215-3247, :3414-3451)
PET counts start at 2: the reference's freqToDistance memo leaves count 1
uninitialised (LooperSolver.hpp:950-951 fill 2..100, :3950 reads 1..100).
Floats are stored as float32 bit patterns (base64, little endian), so the
fixture is exact.  Run where the reference sources exist:
    make -C oracle ref && python tests/golden/make_levels_fixtures.py
"""
import base64
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "oracle", "_ref", "levels_ref")

# Settings of the reference (Settings.hpp:449-461 defaults) and two variants
PARAMS_DEFAULT = dict(count_a=0.5, count_scale=20.0, count_shift=1.0, count_base=0.01,
                      heat_scale=100.0, heat_power=-0.333, heat_scale_inter=100.0,
                      heat_power_inter=-1.0, stretching=2.0,
                      glen_base=0.0, glen_scale=1.0, glen_power=0.5)
PARAMS_POWER = dict(PARAMS_DEFAULT, count_a=-0.333, count_scale=2.5, count_shift=-2.0,
                    count_base=0.05, glen_base=0.1, glen_scale=0.02, glen_power=0.6)
PKEYS = ["count_a", "count_scale", "count_shift", "count_base", "heat_scale", "heat_power",
         "heat_scale_inter", "heat_power_inter", "stretching", "glen_base", "glen_scale",
         "glen_power"]


def f32(a):
    return np.asarray(a, dtype=np.float32)


def b64(a):
    return base64.b64encode(f32(a).ravel().astype("<f4").tobytes()).decode()


def unb64(s, shape=None):
    a = np.frombuffer(base64.b64decode(s), dtype="<f4").copy()
    return a if shape is None else a.reshape(shape)


def fstr(x):
    return repr(float(np.float32(x)))


def param_line(mode, p):
    return mode + " " + " ".join(fstr(p[k]) for k in PKEYS) + "\n"


def tree_text(tree):
    out = [f"{len(tree['chrs'])}\n"]
    for c in tree["chrs"]:
        out.append(f"{c['name']} {len(c['anchors'])} {len(c['arcs'])}\n")
        for (s, e), o in zip(c["anchors"], c["orient"]):
            out.append(f"{s} {e} {o}\n")
        for a, b, sc in c["arcs"]:
            out.append(f"{a} {b} {sc}\n")
    out.append(f"{len(tree['splits'])}\n")
    for chr_name, s, e in tree["splits"]:
        out.append(f"{chr_name} {s} {e}\n")
    return "".join(out)


def run(text):
    r = subprocess.run([BIN], input=text, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, check=True)
    return r.stdout


def hexf(s):
    return float.fromhex(s)


# ---------------------------------------------------------------- inputs
def chromosome(rng, name, n, n_arcs, span, first_free=False):
    gaps = rng.integers(2000, 20000, size=n)
    lens = rng.integers(500, 3000, size=n)
    starts = 10_000 + np.cumsum(gaps + lens) - lens
    anchors = [(int(s), int(s + l)) for s, l in zip(starts, lens)]
    orient = [str(rng.choice(["L", "R", "N"])) for _ in range(n)]
    pairs = set()
    lo = 1 if first_free else 0
    tries = 0
    while len(pairs) < n_arcs and n - lo >= 2 and tries < 50 * (n_arcs + 1):
        tries += 1
        a = int(rng.integers(lo, n - 1))
        b = int(min(n - 1, a + rng.integers(1, span + 1)))
        if a != b:
            pairs.add((a, b))
    arcs = [(a, b, int(rng.integers(2, 60))) for a, b in sorted(pairs)]
    order = rng.permutation(len(arcs))
    arcs = [arcs[i] for i in order]
    return dict(name=name, anchors=anchors, orient=orient, arcs=arcs)


def splits_for(rng, c, k):
    """k predefined segments covering the chromosome (BED order)."""
    lo, hi = c["anchors"][0][0] - 5000, c["anchors"][-1][1] + 5000
    cuts = sorted(int(x) for x in rng.integers(lo + 1, hi, size=k - 1))
    edges = [lo] + cuts + [hi]
    return [(c["name"], edges[i], edges[i + 1]) for i in range(k)]


def tree_cases(rng):
    cases = []
    a = chromosome(rng, "chr1", 30, 18, 3)
    cases.append(dict(name="one_chr_no_splits", chrs=[a], splits=[]))
    b = chromosome(rng, "chr2", 40, 34, 6)
    cases.append(dict(name="one_chr_splits", chrs=[b], splits=splits_for(rng, b, 4)))
    c1 = chromosome(rng, "chrA", 25, 15, 4, first_free=True)
    c2 = chromosome(rng, "chrB", 35, 40, 8)
    cases.append(dict(name="two_chr_splits", chrs=[c1, c2],
                      splits=splits_for(rng, c1, 2) + splits_for(rng, c2, 3)))
    d = chromosome(rng, "chrS", 1, 0, 1)
    cases.append(dict(name="single_anchor", chrs=[d], splits=[]))
    e1 = chromosome(rng, "chrX", 12, 0, 2)
    e2 = chromosome(rng, "chrY", 20, 30, 19)
    cases.append(dict(name="no_arcs_and_dense", chrs=[e1, e2],
                      splits=splits_for(rng, e1, 2) + splits_for(rng, e2, 5)))
    return cases


def parse_tree(out):
    lines = out.splitlines()
    marks = [k for k, l in enumerate(lines) if l.startswith("clusters ")]
    if not marks:   # the reference called error(), which prints and exits
        return dict(error=[l for l in lines if l][-1])
    i = marks[0]
    n = int(lines[i].split()[1])
    clusters = []
    for l in lines[i + 1:i + 1 + n]:
        v = [int(x) for x in l.split()]
        clusters.append(dict(start=v[0], end=v[1], genomic_pos=v[2], base_start=v[3],
                             base_end=v[4], level=v[5], parent=v[6], children=v[8:8 + v[7]]))
    chrs, levels = {}, {}
    for l in lines[i + 1 + n:]:
        t = l.split()
        if t[0] == "chr":
            chrs[t[1]] = dict(first=int(t[2]), root=int(t[3]))
            cur = t[1]
        elif t[0] == "arcs":
            v = [int(x) for x in t[2:]]
            chrs[cur]["arcs"] = [v[k:k + 2] for k in range(0, len(v), 2)]
        elif t[0] == "level":
            levels.setdefault(t[2], {})[t[1]] = [int(x) for x in t[4:]]
    return dict(clusters=clusters, chrs=chrs, levels=levels)


def main():
    if not os.path.exists(BIN):
        raise SystemExit("oracle/_ref/levels_ref not built (make -C oracle ref; needs the reference sources)")
    rng = np.random.default_rng(20261204)
    fx = {"source": "reference LooperSolver host code, oracle/_ref/levels_ref", "params": {
        "default": PARAMS_DEFAULT, "power": PARAMS_POWER}}

    # ---- T1: tree
    trees = tree_cases(rng)
    fx["tree"] = []
    for t in trees:
        out = parse_tree(run(param_line("tree", PARAMS_DEFAULT) + tree_text(t)))
        fx["tree"].append(dict(input=t, output=out))

    # ---- S5: densify (+ dist_to_next / noise of the arc driver)
    fx["densify"] = []
    for ti, ld, fix, pk in [(1, 5, 1, "default"), (1, 3, 0, "power"), (2, 7, 1, "default"),
                            (2, 1, 1, "power"), (0, 2, 1, "default")]:
        t = trees[ti]
        tout = fx["tree"][ti]["output"]
        ibs = [k for k, c in enumerate(tout["clusters"]) if c["level"] == 3]
        ib = max(ibs, key=lambda k: (len(tout["clusters"][k]["children"]), -k))
        if ti == 0:
            ib = min(ibs, key=lambda k: (len(tout["clusters"][k]["children"]), k))
        kids = tout["clusters"][ib]["children"]
        pos = f32(rng.uniform(-5, 5, size=(len(kids), 3)))
        text = (param_line("densify", fx["params"][pk]) + tree_text(t) +
                f"{ib} {ld} {fix}\n" + "".join(" ".join(fstr(x) for x in p) + "\n" for p in pos))
        lines = run(text).splitlines()
        k = next(i for i, l in enumerate(lines) if l.startswith("active "))
        na = int(lines[k].split()[1])
        rows = [l.split() for l in lines[k + 1:k + 1 + na]]
        rest = {l.split()[0]: l.split()[1:] for l in lines[k + 1 + na:]}
        fx["densify"].append(dict(
            tree=ti, ib=ib, loop_density=ld, fix=fix, params=pk, pos=b64(pos),
            active=[int(r[0]) for r in rows], start=[int(r[1]) for r in rows],
            end=[int(r[2]) for r in rows], genomic_pos=[int(r[3]) for r in rows],
            level=[int(r[4]) for r in rows], is_fixed=[int(r[5]) for r in rows],
            out_pos=b64([[hexf(x) for x in r[6:9]] for r in rows]),
            dist_to_next=b64([hexf(x) for x in rest["dist_to_next"]]),
            noise_subanchor=b64([hexf(rest["noise"][0])]), noise_anchor=b64([hexf(rest["noise"][1])]),
            children=[int(x) for x in rest["children"][1:]]))

    # ---- A6: expected distances of one interaction block
    fx["expdist"] = []
    for ti, chr_i, pk, use_heat in [(1, 0, "default", 0), (1, 0, "power", 1), (2, 1, "default", 0),
                                    (2, 0, "power", 0), (4, 1, "default", 1)]:
        t = trees[ti]
        tout = fx["tree"][ti]["output"]
        cname = t["chrs"][chr_i]["name"]
        first = tout["chrs"][cname]["first"]
        anc = t["chrs"][chr_i]["anchors"]
        ibs = [k for k, c in enumerate(tout["clusters"]) if c["level"] == 3
               and first <= c["children"][0] < first + len(anc)]
        ib = max(ibs, key=lambda k: (len(tout["clusters"][k]["children"]), -k))
        n = len(tout["clusters"][ib]["children"])
        longs = []
        for _ in range(12):
            ia, ib2 = rng.integers(0, len(anc), size=2)
            sa = int(rng.integers(anc[ia][0], anc[ia][1] + 1))
            sb = int(rng.integers(anc[ib2][0], anc[ib2][1] + 1))
            longs.append((sa, sb, int(rng.integers(2, 200))))
        longs.append((anc[0][0] - 100, anc[-1][1], 5))                    # outside any anchor
        longs.append((anc[1][0], anc[1][1], 7))                           # same anchor
        heat = f32(rng.uniform(0, 30, size=(n, n)))
        heat = (heat + heat.T) / 2
        text = (param_line("expdist", fx["params"][pk]) + tree_text(t) + f"{cname} {ib} {len(longs)}\n" +
                "".join(f"{a} {b} {s}\n" for a, b, s in longs) + f"{use_heat} 0.5\n")
        if use_heat:
            text += "".join(" ".join(fstr(x) for x in row) + "\n" for row in heat)
        lines = run(text).splitlines()
        k = next(i for i, l in enumerate(lines) if l.startswith("expdist "))
        mat = [[hexf(x) for x in l.split()] for l in lines[k + 1:k + 1 + n]]
        fx["expdist"].append(dict(tree=ti, chr=cname, ib=ib, params=pk, long_arcs=longs,
                                  use_anchor_heatmap=use_heat, influence=0.5,
                                  anchor_heatmap=b64(heat) if use_heat else None,
                                  n=n, out=b64(mat)))

    # ---- H6: distance heatmap
    fx["distheat"] = []
    for n, diag, ind, pk, zero_frac in [(12, 1, 1, "default", 0.1), (40, 2, 0, "default", 0.2),
                                        (64, 1, 1, "default", 0.0), (33, 0, 1, "power", 0.3)]:
        d = np.abs(np.arange(n)[:, None] - np.arange(n)[None, :]).astype(np.float64)
        f = 1.0 / (d ** 3 + 0.1) * rng.uniform(0.5, 1.5, size=(n, n))
        f = f32((f + f.T) / 2)
        f[rng.random((n, n)) < zero_frac] = 0.0
        f[0, n - 1] = 1e-7
        text = (param_line("distheat", fx["params"][pk]) + f"{n} {diag} {ind}\n" +
                "".join(" ".join(fstr(x) for x in row) + "\n" for row in f))
        lines = run(text).splitlines()
        k = next(i for i, l in enumerate(lines) if l.startswith("distheat "))
        hdr = lines[k].split()
        mat = [[hexf(x) for x in l.split()] for l in lines[k + 1:k + 1 + n]]
        fx["distheat"].append(dict(n=n, diag=diag, ind=ind, params=pk, freq=b64(f),
                                   avg_after=b64([hexf(hdr[3])]), out=b64(mat)))

    # ---- energies (pins the oracle's score functions)
    fx["score"] = []
    for n in (8, 24, 50):
        pos = f32(rng.uniform(-3, 3, size=(n, 3)))
        diag = 1
        h = f32(rng.uniform(0.5, 4, size=(n, n)))
        h = (h + h.T) / 2
        h[rng.random((n, n)) < 0.1] = 0.0
        for i in range(n):
            for j in range(n):
                if abs(i - j) < diag:
                    h[i, j] = -1.0
        e = np.full((n, n), -1.0, np.float32)
        for _ in range(n):
            a, b = rng.integers(0, n, size=2)
            if a != b:
                e[a, b] = e[b, a] = rng.uniform(0.3, 3.0)
        e[0, n - 1] = e[n - 1, 0] = 0.0
        np.fill_diagonal(e, 0.0)
        dtn = f32(rng.uniform(0.3, 2.0, size=n))
        text = (param_line("score", PARAMS_DEFAULT) + f"{n}\n" +
                "".join(" ".join(fstr(x) for x in p) + "\n" for p in pos) + f"{diag}\n" +
                "".join(" ".join(fstr(x) for x in row) + "\n" for row in h) +
                "".join(" ".join(fstr(x) for x in row) + "\n" for row in e) +
                " ".join(fstr(x) for x in dtn) + "\n")
        lines = run(text).splitlines()
        sc = next(l for l in lines if l.startswith("score ")).split()[1:]
        cs = next(l for l in lines if l.startswith("constants ")).split()[1:]
        fx["score"].append(dict(n=n, pos=b64(pos), diag=diag, heatmap=b64(h), expdist=b64(e),
                                dist_to_next=b64(dtn), heatmap_score=float(sc[0]),
                                arcs_score=float(sc[1]), smooth_score=float(sc[2]),
                                constants=[hexf(x) for x in cs]))

    path = os.path.join(HERE, "levels_ref.json")
    with open(path, "w") as fh:
        json.dump(fx, fh, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    sys.exit(main())
