"""Host-side rows around the annealing kernels (include/pmmc_levels.h): the
cluster tree and bead indexing, densification, expected distances, the
distance heatmap.  CPU only.

Three layers of checks:
  * the oracle (oracle/levels.py) against tests/golden/levels_ref.json, which
    the reference's own host code wrote (tests/golden/make_levels_fixtures.py);
  * the product (libpmmc_hip.so through pmmc_amd.tree) against the same
    fixture, bit for bit;
  * the product against the oracle on further seeded inputs, plus invariants
    (children partition their parents, genomic order, base ranges, counts).
The fixture also carries the reference's CPU scores, which pin the oracle's
FP64 energies (the reference measures lengths in float: 1e-5 relative).
"""
import copy
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from make_levels_fixtures import chromosome, splits_for, unb64  # noqa: E402

from oracle import binding as O  # noqa: E402
from oracle import levels as L  # noqa: E402
from pmmc_amd import api  # noqa: E402
from pmmc_amd import tree as T  # noqa: E402

FX = json.load(open(os.path.join(ROOT, "tests", "golden", "levels_ref.json")))
PKEYS_MAP = {"count_a": "count_a", "count_scale": "count_scale", "count_shift": "count_shift",
             "count_base": "count_base", "heat_scale": "heat_scale", "heat_power": "heat_power",
             "heat_scale_inter": "heat_scale_inter", "heat_power_inter": "heat_power_inter",
             "stretching": "heat_stretching", "glen_base": "glen_base",
             "glen_scale": "glen_scale", "glen_power": "glen_power"}


def product_params(name_or_dict):
    d = FX["params"][name_or_dict] if isinstance(name_or_dict, str) else name_or_dict
    return T.level_params(**{PKEYS_MAP[k]: v for k, v in d.items()})


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def oracle_tree(case):
    return L.build_tree(case["chrs"], case["splits"])


def product_tree(case):
    return T.Tree(case["chrs"], case["splits"])


def tree_dump_product(t, names):
    cl = [dict(start=c.start, end=c.end, genomic_pos=c.genomic_pos, base_start=c.base_start,
               base_end=c.base_end, level=c.level, parent=c.parent) for c in t.clusters()]
    for i, c in enumerate(cl):
        c["children"] = t.children(i)
    chrs = {}
    for k, name in enumerate(names):
        first, root = t.chromosome(k)
        chrs[name] = dict(first=first, root=root, arcs=t.arcs(k))
    levels = {name: {str(l): t.level(k, l) for l in range(5)} for k, name in enumerate(names)}
    return dict(clusters=cl, chrs=chrs, levels=levels)


def tree_dump_oracle(t):
    cl = [dict(start=c["start"], end=c["end"], genomic_pos=c["genomic_pos"],
               base_start=c["base_start"], base_end=c["base_end"], level=c["level"],
               parent=c["parent"], children=list(c["children"])) for c in t["clusters"]]
    chrs = {n: dict(first=t["chr_first"][n], root=t["chr_root"][n], arcs=t["arcs"][n])
            for n in t["chrs"]}
    levels = {n: {str(l): L.set_level(t, l)[n] for l in range(5)} for n in t["chrs"]}
    return dict(clusters=cl, chrs=chrs, levels=levels)


# ------------------------------------------------------------------ oracle vs reference
@pytest.mark.parametrize("k", range(len(FX["tree"])))
def test_oracle_tree_matches_reference(k):
    case = FX["tree"][k]
    ref = case["output"]
    if "error" in ref:
        with pytest.raises(L.ReferenceError_) as e:
            oracle_tree(case["input"])
        assert str(e.value) == ref["error"]
        return
    assert tree_dump_oracle(oracle_tree(case["input"])) == ref


def _oracle_densified(d):
    t = oracle_tree(FX["tree"][d["tree"]]["input"])
    kids = t["clusters"][d["ib"]]["children"]
    pos = unb64(d["pos"], (len(kids), 3))
    for k, c in enumerate(kids):
        t["clusters"][c]["pos"] = [np.float32(x) for x in pos[k]]
    act = L.densify(t, d["ib"], d["loop_density"], d["fix"])
    return t, act


@pytest.mark.parametrize("k", range(len(FX["densify"])))
def test_oracle_densify_matches_reference(k):
    d = FX["densify"][k]
    t, act = _oracle_densified(d)
    cl = t["clusters"]
    assert act == d["active"] and cl[d["ib"]]["children"] == d["children"]
    assert [cl[c]["start"] for c in act] == d["start"]
    assert [cl[c]["end"] for c in act] == d["end"]
    assert [cl[c]["genomic_pos"] for c in act] == d["genomic_pos"]
    assert [cl[c]["level"] for c in act] == d["level"]
    assert [int(cl[c]["is_fixed"]) for c in act] == d["is_fixed"]
    pos = np.array([[float(x) for x in cl[c]["pos"]] for c in act], np.float32)
    assert np.array_equal(bits(pos), bits(unb64(d["out_pos"], pos.shape)))
    dn, ns, na = L.dist_to_next([cl[c]["genomic_pos"] for c in act], FX["params"][d["params"]])
    assert np.array_equal(bits(dn), bits(unb64(d["dist_to_next"])))
    assert bits(ns) == bits(unb64(d["noise_subanchor"])) and bits(na) == bits(unb64(d["noise_anchor"]))


@pytest.mark.parametrize("k", range(len(FX["expdist"])))
def test_oracle_expected_distances_match_reference(k):
    d = FX["expdist"][k]
    t = oracle_tree(FX["tree"][d["tree"]]["input"])
    heat = unb64(d["anchor_heatmap"], (d["n"], d["n"])) if d["use_anchor_heatmap"] else None
    m = L.expected_distances(t, d["chr"], d["ib"], d["long_arcs"], FX["params"][d["params"]], heat,
                             d["influence"])
    assert np.array_equal(bits(m), bits(unb64(d["out"], (d["n"], d["n"]))))


@pytest.mark.parametrize("k", range(len(FX["distheat"])))
def test_oracle_distance_heatmap_matches_reference(k):
    d = FX["distheat"][k]
    f = unb64(d["freq"], (d["n"], d["n"]))
    h, _ = L.distance_heatmap(f, d["diag"], d["ind"] == 0, FX["params"][d["params"]])
    assert np.array_equal(bits(h), bits(unb64(d["out"], (d["n"], d["n"]))))
    assert bits(L.heat_avg(h)) == bits(unb64(d["avg_after"]))


@pytest.mark.parametrize("k", range(len(FX["score"])))
def test_oracle_energies_match_reference_cpu_scores(k):
    """The oracle's FP64 full energies against the reference's CPU scoring
    (which measures lengths in float)."""
    d = FX["score"][k]
    n = d["n"]
    pos = unb64(d["pos"], (n, 3))
    ks_a, kq_a, ks, kq, ka, wd, wa = d["constants"]
    hm = O.Problem(api.Heatmap(pos=pos, dist=unb64(d["heatmap"], (n, n)), diagonal=d["diag"],
                               skip_below=1e-6))
    ar = O.Problem(api.Arcs(pos=pos, exp_dist=unb64(d["expdist"], (n, n)), k_stretch=ks_a,
                            k_squeeze=kq_a))
    sm = O.Problem(api.Smooth(pos=pos, dist_to_next=unb64(d["dist_to_next"]), k_stretch=ks,
                              k_squeeze=kq, k_angular=ka, w_dist=wd, w_angle=wa))
    assert hm.total() == pytest.approx(d["heatmap_score"], rel=1e-5)
    assert ar.total() == pytest.approx(d["arcs_score"], rel=1e-5)
    assert sm.total() == pytest.approx(d["smooth_score"], rel=1e-5)


# ------------------------------------------------------------------ product vs reference
@pytest.mark.parametrize("k", range(len(FX["tree"])))
def test_product_tree_matches_reference(k):
    case = FX["tree"][k]
    ref = case["output"]
    if "error" in ref:
        with pytest.raises(RuntimeError, match="find split: size mismatch"):
            product_tree(case["input"])
        return
    t = product_tree(case["input"])
    assert tree_dump_product(t, [c["name"] for c in case["input"]["chrs"]]) == ref


@pytest.mark.parametrize("k", range(len(FX["densify"])))
def test_product_densify_matches_reference(k):
    d = FX["densify"][k]
    t = product_tree(FX["tree"][d["tree"]]["input"])
    kids = t.children(d["ib"])
    t.set_positions(kids, unb64(d["pos"], (len(kids), 3)))
    n = t.densify(d["ib"], d["loop_density"], d["fix"])
    act = t.children(d["ib"])
    assert n == len(act) and act == d["active"] == d["children"]
    cl = t.clusters()
    assert [cl[c].start for c in act] == d["start"]
    assert [cl[c].end for c in act] == d["end"]
    assert [cl[c].genomic_pos for c in act] == d["genomic_pos"]
    assert [cl[c].level for c in act] == d["level"]
    assert [cl[c].is_fixed for c in act] == d["is_fixed"]
    assert np.array_equal(bits(t.positions(act)), bits(unb64(d["out_pos"], (len(act), 3))))
    p = product_params(d["params"])
    na, ns = t.set_dist_to_next(d["ib"], p)
    dn = np.array([t.clusters()[c].dist_to_next for c in act[:-1]], np.float32)
    assert np.array_equal(bits(dn), bits(unb64(d["dist_to_next"])))
    assert bits(ns) == bits(unb64(d["noise_subanchor"])) and bits(na) == bits(unb64(d["noise_anchor"]))


@pytest.mark.parametrize("k", range(len(FX["expdist"])))
def test_product_expected_distances_match_reference(k):
    d = FX["expdist"][k]
    case = copy.deepcopy(FX["tree"][d["tree"]]["input"])
    ci = [c["name"] for c in case["chrs"]].index(d["chr"])
    case["chrs"][ci]["long_arcs"] = d["long_arcs"]
    t = product_tree(case)
    p = product_params(d["params"])
    p.use_anchor_heatmap = d["use_anchor_heatmap"]
    p.anchor_heatmap_influence = d["influence"]
    heat = unb64(d["anchor_heatmap"], (d["n"], d["n"])) if d["use_anchor_heatmap"] else None
    m = t.expected_distances(ci, d["ib"], p, heat)
    assert np.array_equal(bits(m), bits(unb64(d["out"], (d["n"], d["n"]))))


@pytest.mark.parametrize("k", range(len(FX["distheat"])))
def test_product_distance_heatmap_matches_reference(k):
    d = FX["distheat"][k]
    f = unb64(d["freq"], (d["n"], d["n"]))
    h, _ = T.distance_heatmap(f, d["diag"], d["ind"] == 0, product_params(d["params"]))
    assert np.array_equal(bits(h), bits(unb64(d["out"], (d["n"], d["n"]))))


# ------------------------------------------------------------------ product vs oracle, wider
def random_case(seed):
    rng = np.random.default_rng(seed)
    n_chr = int(rng.integers(1, 4))
    chrs, splits = [], []
    for c in range(n_chr):
        n = int(rng.integers(2, 60))
        ch = chromosome(rng, f"c{c}", n, int(rng.integers(0, 2 * n)), int(rng.integers(1, 12)),
                        first_free=bool(rng.integers(0, 2)))
        chrs.append(ch)
        splits += splits_for(rng, ch, int(rng.integers(1, 5)))
    if seed % 3 == 0:
        splits = []
    return dict(chrs=chrs, splits=splits), rng


@pytest.mark.parametrize("seed", range(12))
def test_product_matches_oracle_on_random_genomes(seed):
    case, rng = random_case(seed)
    names = [c["name"] for c in case["chrs"]]
    try:
        ot = oracle_tree(case)
    except L.ReferenceError_:
        with pytest.raises(RuntimeError):
            product_tree(case)
        return
    pt = product_tree(case)
    assert tree_dump_product(pt, names) == tree_dump_oracle(ot)
    pkey = ["default", "power"][seed % 2]
    p = product_params(pkey)
    # expected distances and densification of every interaction block
    for ci, name in enumerate(names):
        for ib in pt.level(ci, 2):
            m = pt.expected_distances(ci, ib, p)
            mo = L.expected_distances(ot, name, ib, [], FX["params"][pkey])
            assert np.array_equal(bits(m), bits(mo)), (name, ib)
        for ib in pt.level(ci, 2)[:3]:
            kids = pt.children(ib)
            pos = rng.uniform(-4, 4, size=(len(kids), 3)).astype(np.float32)
            pt.set_positions(kids, pos)
            for k, c in enumerate(kids):
                ot["clusters"][c]["pos"] = [np.float32(x) for x in pos[k]]
            ld = int(rng.integers(0, 8))
            pt.densify(ib, ld, True)
            L.densify(ot, ib, ld, True)
            act = pt.children(ib)
            assert act == ot["clusters"][ib]["children"]
            assert np.array_equal(bits(pt.positions(act)),
                                  bits([[float(x) for x in ot["clusters"][c]["pos"]] for c in act]))


@pytest.mark.parametrize("seed", range(6))
def test_product_distance_heatmap_matches_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 70))
    f = rng.gamma(0.5, 2.0, size=(n, n)).astype(np.float32)
    f[rng.random((n, n)) < 0.2] = 0.0
    diag = int(rng.integers(0, 3))
    for inter in (False, True):
        pkey = ["default", "power"][seed % 2]
        h, avg = T.distance_heatmap(f, diag, inter, product_params(pkey))
        ho, avgo = L.distance_heatmap(f, diag, inter, FX["params"][pkey])
        assert np.array_equal(bits(h), bits(ho))
        assert bits(avg) == bits(avgo) or (np.isnan(avg) and np.isnan(avgo))


def test_freq_to_distance_matches_oracle_and_branches():
    for pkey in ("default", "power"):
        p = product_params(pkey)
        for f in list(range(0, 130)) + [1000, 123456]:
            assert bits(T.freq_to_distance(f, p)) == bits(L.freq_to_distance(f, FX["params"][pkey]))
    # power-law branch: base + scale * max(1, f + shift)^a (LooperSolver.hpp:3955-3961)
    p = T.level_params(count_a=-0.5, count_scale=4.0, count_shift=0.0, count_base=1.0)
    assert T.freq_to_distance(16, p) == np.float32(1.0 + 4.0 / 4.0)
    assert T.freq_to_distance(0, p) == np.float32(1.0 + 4.0)       # clamped to 1
    # exponential branch: base + scale / exp(a * (f + shift))
    p = T.level_params(count_a=1.0, count_scale=2.0, count_shift=0.0, count_base=0.0)
    assert T.freq_to_distance(0, p) == np.float32(2.0)
    assert T.genomic_length_to_distance(4000, T.level_params()) == pytest.approx(2.0, abs=0)


# ------------------------------------------------------------------ invariants and errors
@pytest.mark.parametrize("seed", range(6))
def test_tree_invariants(seed):
    case, _ = random_case(1000 + seed)
    try:
        t = product_tree(case)
    except RuntimeError:
        return
    cl = t.clusters()
    for ci, ch in enumerate(case["chrs"]):
        first, root = t.chromosome(ci)
        na = len(ch["anchors"])
        assert cl[root].level == 1 and cl[root].parent == -1
        anchors = t.level(ci, 3)
        assert anchors == list(range(first, first + na))   # every anchor once, in order
        for lvl in (1, 2, 3):
            for c in t.level(ci, lvl):
                kids = t.children(c)
                if cl[c].level == 4:
                    continue
                assert kids and all(cl[k].parent == c for k in kids)
                assert cl[c].start == cl[kids[0]].start and cl[c].end == cl[kids[-1]].end
                assert [cl[k].genomic_pos for k in kids] == sorted(cl[k].genomic_pos for k in kids)
                if cl[c].level == 3:
                    assert kids == list(range(cl[c].base_start, cl[c].base_end + 1))
                else:
                    assert cl[c].base_start == cl[kids[0]].base_start
                    assert cl[c].base_end == cl[kids[-1]].base_end
        # every arc is shifted to global indices
        loc = {(a + first, b + first) for a, b, _ in ch["arcs"]}
        assert {tuple(x) for x in t.arcs(ci)} == loc


def test_densify_counts_and_spacing():
    case = dict(chrs=[dict(name="x", anchors=[(0, 100), (1100, 1200), (5200, 5300)],
                           arcs=[(0, 2, 5)])], splits=[])
    t = product_tree(case)
    ib = t.level(0, 2)[0]
    assert t.children(ib) == [0, 1, 2]
    t.set_positions([0, 1, 2], [[0, 0, 0], [3, 0, 0], [3, 4, 0]])
    assert t.densify(ib, 4, True) == 3 + 2 * 4
    act = t.children(ib)
    cl = t.clusters()
    # 1000 bp between anchors 0 and 1 -> steps of 200 bp from the left end
    assert [cl[c].genomic_pos for c in act[1:5]] == [300, 500, 700, 900]
    assert np.allclose(t.positions(act[1:5])[:, 0], [0.6, 1.2, 1.8, 2.4])
    assert [cl[c].is_fixed for c in act] == [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]
    assert all(cl[c].level == 5 and cl[c].parent == -1 for c in act if c > 2)


def test_tree_rejects_bad_input():
    with pytest.raises(RuntimeError, match="no anchors"):
        T.Tree([dict(name="a", anchors=[], arcs=[])])
    with pytest.raises(RuntimeError, match="out of range"):
        T.Tree([dict(name="a", anchors=[(0, 1), (5, 6)], arcs=[(0, 2, 3)])])
    with pytest.raises(RuntimeError, match="no segment split found for b"):
        T.Tree([dict(name="a", anchors=[(0, 1), (5, 6)], arcs=[]),
                dict(name="b", anchors=[(0, 1), (5, 6)], arcs=[])], splits=[("a", 0, 10)])
    # a later single-anchor chromosome under predefined segments: the reference
    # reads children[0] of an empty list; here it is an error
    with pytest.raises(RuntimeError, match="no segment"):
        T.Tree([dict(name="a", anchors=[(0, 1), (5, 6)], arcs=[]),
                dict(name="b", anchors=[(0, 1)], arcs=[])], splits=[("a", 0, 10), ("b", 0, 10)])
