"""pmmc_singleton_heatmap on the GPU against the reference's own
createSingletonHeatmap output (tests/golden/singleton_ref.json) and, at
millions of records, against the oracle's restatement: bit-exact (NaN where
both divide 0 by a zero bin length), same counts.  Covers the LDS-histogram
path (few bins), the device-matrix path (many bins) and unsorted breaks (the
linear first-match scan)."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from make_levels_fixtures import unb64  # noqa: E402

from oracle import levels as L  # noqa: E402
from pmmc_amd.api import PmmcError  # noqa: E402

pytestmark = pytest.mark.gpu

FX = json.load(open(os.path.join(ROOT, "tests", "golden", "singleton_ref.json")))


def same_bits(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("k", range(len(FX["singleton"])))
def test_singleton_heatmap_matches_reference(ctx, k):
    c = FX["singleton"][k]
    print(f"singleton fixture {k}: n={c['n']} records={len(c['records'])}", flush=True)
    heat, cnt = ctx.singleton_heatmap(c["chrs"], c["clusters"], [tuple(r) for r in c["records"]],
                                      tuple(c["region"]), [tuple(x) for x in c["centromeres"]],
                                      [tuple(x) for x in c["long_arcs"]], c["long_scale"],
                                      c["long_power"], c["diag"])
    assert same_bits(heat, unb64(c["out"], (c["n"], c["n"])))
    _, ocnt = L.singleton_heatmap(c["chrs"], c["clusters"], [tuple(r) for r in c["records"]],
                                  tuple(c["region"]), [tuple(x) for x in c["centromeres"]],
                                  [tuple(x) for x in c["long_arcs"]], c["long_scale"],
                                  c["long_power"], c["diag"])
    assert {k2: cnt[k2] for k2 in ocnt} == ocnt


def random_level(rng, n_chr, per_chr, shuffle=False):
    chrs, clusters = [], []
    for c in range(n_chr):
        gaps = rng.integers(2000, 30000, size=per_chr)
        lens = rng.integers(500, 8000, size=per_chr)
        starts = 50_000 + np.cumsum(gaps + lens) - lens
        cl = [(int(s), int(s + l)) for s, l in zip(starts, lens)]
        if shuffle:
            rng.shuffle(cl)
        chrs.append(f"c{c}")
        clusters.append(cl)
    return chrs, clusters


def random_records(rng, chrs, clusters, n):
    k = len(chrs)
    hi = np.array([max(e for _, e in cl) + 20_000 for cl in clusters], np.int64)
    c1 = rng.integers(0, k, size=n)
    c2 = np.where(rng.random(n) < 0.7, c1, rng.integers(0, k, size=n))
    s1 = (rng.random(n) * hi[c1]).astype(np.int64) - 100
    s2 = (rng.random(n) * hi[c2]).astype(np.int64) - 100
    rec = dict(chr1=c1, start1=s1, end1=s1 + rng.integers(20, 400, size=n), chr2=c2, start2=s2,
               end2=s2 + rng.integers(20, 400, size=n), score=rng.integers(1, 20, size=n))
    rec["chr1"] = np.where(rng.random(n) < 0.01, -1, rec["chr1"])      # not processed
    return {key: v.astype(np.int32) for key, v in rec.items()}


def oracle_from_arrays(chrs, clusters, rec, **kw):
    names = list(chrs) + ["chrUn"]
    tup = list(zip([names[i] for i in rec["chr1"]], rec["start1"].tolist(), rec["end1"].tolist(),
                   [names[i] for i in rec["chr2"]], rec["start2"].tolist(), rec["end2"].tolist(),
                   rec["score"].tolist()))
    return L.singleton_heatmap(chrs, clusters, tup, **kw)


@pytest.mark.parametrize("n_chr,per_chr,shuffle,n_rec", [
    (3, 4, False, 2_000_000),     # 12 bins: LDS histogram
    (3, 60, False, 2_000_000),    # 180 bins: device matrix
    (2, 30, True, 500_000)])      # unsorted breaks: linear first-match scan
def test_singleton_heatmap_matches_oracle_at_scale(ctx, n_chr, per_chr, shuffle, n_rec):
    rng = np.random.default_rng(n_chr * 1000 + per_chr)
    chrs, clusters = random_level(rng, n_chr, per_chr, shuffle)
    rec = random_records(rng, chrs, clusters, n_rec)
    longs = [(c, int(cl[0][0]) + 10, int(cl[-1][1]) - 10, 7) for c, cl in zip(chrs, clusters)]
    cents = [(chrs[0], int(clusters[0][1][0]), int(clusters[0][2][1]))]
    kw = dict(region=(0, 0), centromeres=cents, long_arcs=longs, long_scale=10.0, long_power=2.0,
              diag=1)
    print(f"singleton scale: chr={n_chr} bins={n_chr * per_chr} shuffle={shuffle} records={n_rec}",
          flush=True)
    heat, cnt = ctx.singleton_heatmap(chrs, clusters, rec, **kw)
    want, ocnt = oracle_from_arrays(chrs, clusters, rec, **kw)
    assert same_bits(heat, want)
    assert {k: cnt[k] for k in ocnt} == ocnt


def test_singleton_heatmap_errors(ctx):
    chrs, clusters = ["a"], [[(1000, 2000), (5000, 6000)]]
    with pytest.raises(PmmcError):   # a position past the last break (the reference exits)
        ctx.singleton_heatmap(chrs, clusters, [("a", 1_500_000_000, 1_500_000_100, "a", 10, 20, 1)])
    with pytest.raises(PmmcError):   # a long arc past the last break
        ctx.singleton_heatmap(chrs, clusters, [("a", 10, 20, "a", 5500, 5600, 1)],
                              long_arcs=[("a", 10, 1_500_000_000, 3)])
    heat, cnt = ctx.singleton_heatmap(chrs, clusters, [])   # no records
    assert cnt["added"] == 0 and heat.shape == (2, 2)
