"""Heatmap input construction upstream of createDistanceHeatmap, on the CPU:
the oracle's restatement of LooperSolver::createSingletonHeatmap
(LooperSolver.hpp:2517-2749) and of the three normalisations (:2996-3040,
:3146-3172, :3106-3144), and the product's host normalisations
(pmmc_normalize_heatmap*), against the reference's own output for the same
inputs (tests/golden/singleton_ref.json, written by
tests/golden/make_singleton_fixtures.py with oracle/_ref/levels_ref).
Bit-exact, NaN where the reference divides 0 by a zero bin length."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from make_levels_fixtures import unb64  # noqa: E402

from oracle import levels as L  # noqa: E402
from pmmc_amd import tree as T  # noqa: E402

FX = json.load(open(os.path.join(ROOT, "tests", "golden", "singleton_ref.json")))


def same_bits(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))


def oracle_case(c):
    return L.singleton_heatmap(c["chrs"], c["clusters"], [tuple(r) for r in c["records"]],
                               tuple(c["region"]), [tuple(x) for x in c["centromeres"]],
                               [tuple(x) for x in c["long_arcs"]], c["long_scale"], c["long_power"],
                               c["diag"])


@pytest.mark.parametrize("k", range(len(FX["singleton"])))
def test_oracle_singleton_heatmap_matches_reference(k):
    c = FX["singleton"][k]
    heat, cnt = oracle_case(c)
    assert same_bits(heat, unb64(c["out"], (c["n"], c["n"])))
    assert cnt["added"] > 0 and cnt["diagonal"] > 0 and cnt["bad_chr"] > 0


@pytest.mark.parametrize("k", range(len(FX["normalize"])))
def test_normalize_matches_reference(k):
    c = FX["normalize"][k]
    n = sum(c["sizes"])
    h = unb64(c["heat"], (n, n))
    want = {key: unb64(c[key], (n, n)) for key in ("normalized", "diagonal_total", "inter")}
    # oracle
    assert same_bits(L.normalize_heatmap(h), want["normalized"])
    assert same_bits(L.normalize_heatmap_diagonal_total(h, c["val"]), want["diagonal_total"])
    assert same_bits(L.normalize_heatmap_inter(h, c["sizes"], c["scale"]), want["inter"])
    # product (host C++)
    assert same_bits(T.normalize_heatmap(h), want["normalized"])
    assert same_bits(T.normalize_heatmap_diagonal_total(h, c["val"]), want["diagonal_total"])
    assert same_bits(T.normalize_heatmap_inter(h, c["sizes"], c["scale"]), want["inter"])


def test_singleton_closed_form():
    # two chromosomes of two clusters each: breaks (0, mid, 1e9); four records
    # with hand-computed bins, one diagonal, one unknown chromosome
    chrs = ["a", "b"]
    clusters = [[(1000, 2000), (5000, 6000)], [(100, 300), (900, 1000)]]  # mids 3500, 600
    recs = [("a", 1200, 1300, "a", 5200, 5300, 3),      # bins 0 -> 1
            ("a", 3500, 3600, "b", 700, 800, 2),        # 3500 is on the break: bin 0 -> 3
            ("b", 100, 150, "b", 200, 250, 5),          # same bin: diagonal
            ("x", 1, 2, "a", 3, 4, 9)]                  # unknown chromosome
    heat, cnt = L.singleton_heatmap(chrs, clusters, recs, diag=1)
    assert cnt == dict(added=2, diagonal=1, bad_chr=1)
    # lengths end - start + 1 between breaks whose outer ends became the lowest
    # start / highest end of the added records: a: (1200, 3500, 5300) -> 2301,
    # 1801 bp; b: (700, 600, 800) -> -99 (the lowest start lies past the
    # midpoint), 201 bp
    la = np.float32([2301, 1801, -99, 201]) / np.float32(1e6)
    raw = np.zeros((4, 4), np.float32)
    raw[0, 1] = 3
    raw[0, 3] = 2
    with np.errstate(divide="ignore", invalid="ignore"):
        q = raw / np.outer(la, la)   # 0 / a negative length is -0, as in C
    iu = np.triu_indices(4, 1)
    want = np.zeros((4, 4), np.float32)
    want[iu] = q[iu]
    want[(iu[1], iu[0])] = q[iu]
    assert same_bits(heat, want)


def test_centromere_length_cases():
    # the four overlap cases of getGenomicLengthExcludingCentromere (:5916-5935)
    f = L.length_excluding_centromere
    assert f(None, 10, 20) == 11
    assert f((30, 40), 10, 20) == 11 and f((1, 5), 10, 20) == 11    # disjoint
    assert f((5, 15), 10, 20) == 6                                  # covers the start
    assert f((5, 25), 10, 20) == 0                                  # covers the bin
    assert f((12, 15), 10, 20) == 7                                 # inside
    assert f((15, 25), 10, 20) == 5                                 # covers the end
