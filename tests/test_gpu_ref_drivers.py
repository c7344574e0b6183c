"""The reference's own level drivers against the product's.

oracle/_ref/levels_ref is the reference's LooperSolver compiled in place
(oracle/Makefile).  Its drive_arcs / drive_heatmap modes run the reference's
reconstructClusterArcsDistances (LooperSolver.hpp:4303-4487, in the block loop
of :4027-4138) and reconstructClustersHeatmapSingleLevel (:1258-1422); their
annealing calls go through the product's reference-side glue
(pmmc_amd/integration/reference_glue.cpp) to libpmmc_hip.so on the GPU, with
the glue's per-call trace (PMMC_GLUE_TRACE) on.  Three checks per driver:

* the glue: every traced call, repeated as a direct C-ABI call with the same
  problem, schedule preset, seed and segment, gives the positions the glue
  wrote back into clusters[] (non-fixed beads only at the smooth level), the
  value the member returned and its total_mc_steps increment, bit for bit;
* with the restart noise set to 0 the two drivers are deterministic: the
  product's pmmc_arcs_level / pmmc_heatmap_level (segments numbered in the
  reference's call order, the same chains per call) leave the same active
  regions, densified beads, dist_to_next, fixed flags and positions as the
  reference's drivers, bit for bit (the anchor stage anneals with its step
  > 0; at the sub-anchor and heatmap levels a zero noise is a zero step);
* with the restart noise on (glibc rand() there, Philox here) the drivers
  agree statistically over 16 seeds: the summed best scores of the two
  16-sample sets have medians within the larger interquartile range, and a
  Mann-Whitney test does not separate them (p > 1e-3).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest
from scipy import stats as sps

from oracle import refdriver
from pmmc_amd import api
from pmmc_amd import tree as T

from test_gpu_levels import PARAMS, genome, product_params

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not refdriver.available(),
                                 reason="oracle/_ref/levels_ref not built (reference sources absent)")]

HEADER_KEYS = ["count_a", "count_scale", "count_shift", "count_base", "heat_scale", "heat_power",
               "heat_scale_inter", "heat_power_inter", "stretching", "glen_base", "glen_scale",
               "glen_power"]
F = np.float32


def f32(x):
    return float(F(x))


# [simulation_arcs] / [simulation_arcs_smooth] / [simulation_heatmap] values,
# float32 as the reference's Settings hold them
ARCS = dict(max_temp=f32(5.0), dt_temp=f32(0.9995), tjs=f32(50.0), tjc=f32(20.0),
            improvement=f32(0.99), min_successes=5, steps=1536)
SMOOTH = dict(ARCS, steps=1024)
HEAT = dict(max_temp=f32(5.0), dt_temp=f32(0.9999), tjs=f32(50.0), tjc=f32(20.0),
            improvement=f32(0.99), min_successes=5, steps=2048)
SPRINGS = dict(ks_arcs=1.0, kq_arcs=1.0, ks=0.1, kq=0.1, ka=0.1, wd=1.0, wa=1.0)
FAILS = 3
CHAINS = 256


def lvl_settings(d):
    return api.settings(max_temp=d["max_temp"], dt_temp=d["dt_temp"], improvement=d["improvement"],
                        tjs=d["tjs"], tjc=d["tjc"], min_successes=d["min_successes"],
                        steps=d["steps"], milestone_fails=FAILS)


def sline(d):
    return (f"{d['max_temp']!r} {d['dt_temp']!r} {d['tjs']!r} {d['tjc']!r} {d['improvement']!r} "
            f"{d['min_successes']} {d['steps']}")


def tree_text(case):
    out = [f"{len(case['chrs'])}"]
    for c in case["chrs"]:
        out.append(f"{c['name']} {len(c['anchors'])} {len(c['arcs'])}")
        out += [f"{s} {e} N" for s, e in c["anchors"]]
        out += [f"{a} {b} {sc}" for a, b, sc in c["arcs"]]
    out.append(str(len(case["splits"])))
    out += [f"{n} {s} {e}" for n, s, e in case["splits"]]
    return "\n".join(out) + "\n"


def header(mode, p):
    return mode + " " + " ".join(repr(f32(p[k])) for k in HEADER_KEYS) + "\n"


def run_ref(text):
    """Run levels_ref with the glue trace on; returns (stdout after @@drive,
    trace records)."""
    with tempfile.TemporaryDirectory() as d:
        tr = os.path.join(d, "trace.txt")
        env = dict(os.environ, PMMC_GLUE_TRACE=tr)
        r = subprocess.run([refdriver.REF], input=text, capture_output=True, text=True, env=env,
                           cwd=d, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        calls = parse_trace(open(tr).read()) if os.path.exists(tr) else []
    lines = r.stdout.splitlines()
    k = [i for i, ln in enumerate(lines) if ln.startswith("@@drive")][0]
    return lines[k:], calls


def parse_trace(txt):
    calls, cur = [], None
    for ln in txt.splitlines():
        t = ln.split()
        if not t:
            continue
        if t[0] == "@@call":
            cur = dict(level=int(t[1]), seed=int(t[2]), segment=int(t[3]), chains=int(t[4]),
                       n=int(t[5]), step=float.fromhex(t[6]), steps_before=int(t[7]))
            calls.append(cur)
        elif t[0] in ("pos", "mat", "loopk", "out"):
            cur[t[0]] = np.array([float.fromhex(x) for x in t[2:]], np.float32)
        elif t[0] in ("fixed", "loops", "bnd"):
            cur[t[0]] = np.array([int(x) for x in t[2:]], np.int32)
        elif t[0] == "diag":
            cur["diag"] = int(t[1])
        elif t[0] == "ret":
            cur["ret"] = float.fromhex(t[1])
            cur["steps_after"] = int(t[2])
    return calls


def parse_drive(lines):
    _, steps, nb = lines[0].split()
    ids = [int(x) for x in lines[1].split()]
    nc = int(lines[2].split()[1])
    cl = []
    for ln in lines[3:3 + nc]:
        t = ln.split()
        cl.append(dict(pos=np.array([float.fromhex(x) for x in t[:3]], np.float32),
                       fixed=int(t[3]), dtn=float.fromhex(t[4]),
                       children=[int(x) for x in t[6:6 + int(t[5])]]))
    return int(steps), ids, cl


def check_glue_calls(ctx, calls):
    """Each traced glue call, repeated through the C ABI directly."""
    assert calls
    for c in calls:
        n = c["n"]
        pos0 = c["pos"].reshape(n, 3)
        fixed = c["fixed"].astype(np.uint8)
        if c["level"] == 1:
            prob = api.Heatmap(pos=pos0, dist=c["mat"].reshape(n, n), fixed=fixed,
                               diagonal=c["diag"], skip_below=1e-3,
                               boundaries=c["bnd"] if c["bnd"].size else None)
            sched = api.schedule_preset(1, 1, lvl_settings(HEAT), c["step"])
        elif c["level"] == 2:
            lp = c["loops"].reshape(-1, 2)
            prob = api.Arcs(pos=pos0, exp_dist=c["mat"].reshape(n, n), fixed=fixed,
                            loop_pairs=lp if lp.size else None,
                            loop_params=c["loopk"].reshape(-1, 2) if lp.size else None,
                            k_stretch=SPRINGS["ks_arcs"], k_squeeze=SPRINGS["kq_arcs"])
            sched = api.schedule_preset(2, 1, lvl_settings(ARCS), c["step"])
        else:
            prob = api.Smooth(pos=pos0, dist_to_next=c["mat"], fixed=fixed, k_stretch=SPRINGS["ks"],
                              k_squeeze=SPRINGS["kq"], k_angular=SPRINGS["ka"],
                              w_dist=SPRINGS["wd"], w_angle=SPRINGS["wa"])
            sched = api.schedule_preset(3, 1, lvl_settings(SMOOTH), c["step"])
        out, st = ctx.anneal(prob, sched, seed=c["seed"], segment=c["segment"],
                             n_chains=c["chains"])
        want = c["out"].reshape(n, 3)
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32)), (c["level"], c["segment"])
        if c["level"] == 3:       # fixed beads keep the caller's positions
            assert np.array_equal(want[fixed == 1], pos0[fixed == 1])
        ret = st["energy"] if c["level"] == 2 else st["energy_resummed"]
        assert F(ret) == F(c["ret"]), (c["level"], c["segment"], ret, c["ret"])
        assert c["steps_after"] - c["steps_before"] == st["best_chain_moves"]


# ---------------------------------------------------------------- arc level
def arcs_case(seed):
    case, rng = genome(seed, n_chr=1, n_anchors=(24,), span=4, splits_per_chr=3)
    return case, rng


def ref_drive_arcs(case, params, seed, segment, ib_pos, chrom=None, glob=None):
    """glob: (stress iterations, stress lr) to run the global refinement, with
    the case's long arcs; chrom "*" drives every chromosome."""
    p = params
    text = header("drive_arcs", p) + tree_text(case)
    text += f"{chrom or case['chrs'][0]['name']}\n"
    text += (f"{p['loop_density']} {f32(p['noise_anchor'])!r} {f32(p['noise_subanchor'])!r} "
             f"{f32(p['noise_anchor_small'])!r} {p['restarts_anchor']} {p['restarts_subanchor']} "
             f"{p['use2d']} {FAILS}\n")
    text += sline(ARCS) + "\n" + sline(SMOOTH) + "\n"
    text += " ".join(repr(f32(SPRINGS[k])) for k in ("ks_arcs", "kq_arcs", "ks", "kq", "ka", "wd", "wa"))
    text += f"\n{seed} {segment} {CHAINS}\n"
    text += " ".join(float(x).hex() for x in np.asarray(ib_pos, np.float32).reshape(-1)) + "\n"
    if glob is not None:
        longs = [(c["name"], s0, e0, sc) for c in case["chrs"] for s0, e0, sc in c.get("long_arcs", [])]
        text += f"global 1 {glob[0]} {f32(glob[1])!r} {len(longs)}\n"
        text += "".join(f"{n} {s0} {e0} {sc}\n" for n, s0, e0, sc in longs)
    lines, calls = run_ref(text)
    steps, ibs, cl = parse_drive(lines)
    return steps, ibs, cl, calls


def product_drive_arcs(ctx, case, params, seed, segment, ib_pos, all_chrs=False, glob=None):
    pt = T.Tree(case["chrs"], case["splits"])
    chrs = range(len(case["chrs"])) if all_chrs else [0]
    ibs = [ib for c in chrs for ib in pt.level(c, 2)]
    pt.set_positions(ibs, np.asarray(ib_pos, np.float32))
    po = T.arcs_level_opts(arcs=lvl_settings(ARCS), smooth=lvl_settings(SMOOTH),
                           chains_per_restart=CHAINS, k_stretch_arcs=SPRINGS["ks_arcs"],
                           k_squeeze_arcs=SPRINGS["kq_arcs"], k_stretch=SPRINGS["ks"],
                           k_squeeze=SPRINGS["kq"], k_angular=SPRINGS["ka"], w_dist=SPRINGS["wd"],
                           w_angle=SPRINGS["wa"])
    if glob is not None:
        po.global_refine = 1
        po.stress.max_iter = glob[0]
        po.stress.lr = glob[1]
    rep = T.arcs_level(ctx, pt, -1 if all_chrs else 0, product_params(params), po, seed=seed,
                       segment=segment)
    return pt, ibs, rep


def block_scores(calls, level):
    """Best-of per block from the glue's returned values (first on ties),
    summed: the score the reference's driver kept.  The calls of one block's
    stage are consecutive, of one level and one bead count."""
    groups, prev = [], None
    for c in calls:
        if c["level"] != level:
            prev = None
            continue
        if prev is None or prev["n"] != c["n"]:
            groups.append([c])
        else:
            groups[-1].append(c)
        prev = c
    total = 0.0
    for g in groups:
        b = -1.0
        for c in g:
            if c["ret"] < b or b < 0:
                b = c["ret"]
        total += b
    return total


@pytest.mark.parametrize("seed", [3, 8])
def test_ref_arcs_driver_noise_free_bit_exact(ctx, seed):
    case, rng = arcs_case(seed)
    params = dict(PARAMS, noise_anchor_small=0.0, noise_subanchor=0.0, restarts_anchor=2,
                  restarts_subanchor=2, loop_density=2)
    pt0 = T.Tree(case["chrs"], case["splits"])
    ib_pos = np.cumsum(rng.normal(0, 1.0, size=(len(pt0.level(0, 2)), 3)), 0).astype(np.float32)
    print(f"reference arc driver seed={seed} blocks={len(ib_pos)}", flush=True)
    steps, ibs, rcl, calls = ref_drive_arcs(case, params, 77, 500, ib_pos)
    pt, pibs, rep = product_drive_arcs(ctx, case, params, 77, 500, ib_pos)
    assert ibs == pibs
    assert len(pt) == len(rcl)                       # the same beads after densification
    pcl = pt.clusters()
    for i, (a, b) in enumerate(zip(pcl, rcl)):
        assert bool(a.is_fixed) == bool(b["fixed"]), i
        assert np.float32(a.dist_to_next) == np.float32(b["dtn"]), i
    for ib in ibs:
        assert pt.children(ib) == rcl[ib]["children"]   # active regions, densified order
    got = pt.positions(list(range(len(pt))))
    want = np.array([c["pos"] for c in rcl], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the anchor stage really annealed (a zero restart noise, not a zero step):
    # the anchors of a block no longer sit at the block's position
    moved = [any(not np.array_equal(want[c], ib_pos[k]) for c in rcl[ib]["children"] if rcl[c]["fixed"])
             for k, ib in enumerate(ibs) if len(rcl[ib]["children"]) > 1]
    assert moved and all(moved)
    assert rep["score_anchor"] == pytest.approx(block_scores(calls, 2), rel=1e-6)
    check_glue_calls(ctx, calls)


def global_case(seed):
    """Two chromosomes over the same genomic range, long arcs, counts 2-6."""
    case, rng = genome(seed, n_chr=2, n_anchors=(24, 20), span=4, splits_per_chr=2)
    for c in case["chrs"]:
        c["arcs"] = [(a, b, int(rng.integers(2, 7))) for a, b, _ in c["arcs"]]
        anc, longs = c["anchors"], []
        for _ in range(6):
            a, b = sorted(int(x) for x in rng.integers(0, len(anc), size=2))
            longs.append((int(rng.integers(anc[a][0], anc[a][1] + 1)),
                          int(rng.integers(anc[b][0], anc[b][1] + 1)), int(rng.integers(2, 7))))
        c["long_arcs"] = longs
    return case, rng


@pytest.mark.parametrize("stress_iters", [0, 600])
def test_ref_arcs_driver_global_refinement_bit_exact(ctx, stress_iters):
    """The reference's driver with useGlobalAnchorRefinement over two
    chromosomes, noise-free: per chromosome its blocks, then the global pass
    (calcAnchorExpectedDistancesHeatmap and reconstructClusterArcsDistances
    over all anchors through the glue, stressRefineAnchors, the sub-anchor
    re-placement), against pmmc_arcs_level with global_refine: every cluster's
    position, fixed flag and dist_to_next, bit for bit."""
    case, rng = global_case(17)
    params = dict(PARAMS, noise_anchor_small=0.0, noise_subanchor=0.0, restarts_anchor=2,
                  restarts_subanchor=2, loop_density=2)
    pt0 = T.Tree(case["chrs"], case["splits"])
    n_ib = sum(len(pt0.level(c, 2)) for c in range(2))
    ib_pos = np.cumsum(rng.normal(0, 1.0, size=(n_ib, 3)), 0).astype(np.float32)
    print(f"reference global refinement stress_iters={stress_iters}", flush=True)
    _, ibs, rcl, calls = ref_drive_arcs(case, params, 41, 300, ib_pos, chrom="*",
                                        glob=(stress_iters, 0.05))
    pt, pibs, rep = product_drive_arcs(ctx, case, params, 41, 300, ib_pos, all_chrs=True,
                                       glob=(stress_iters, 0.05))
    assert ibs == pibs and len(pt) == len(rcl)
    pcl = pt.clusters()
    for i, (a, b) in enumerate(zip(pcl, rcl)):
        assert bool(a.is_fixed) == bool(b["fixed"]), i
        assert np.float32(a.dist_to_next) == np.float32(b["dtn"]), i
    got = pt.positions(list(range(len(pt))))
    want = np.array([c["pos"] for c in rcl], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert rep["reinterpolated"] > 0
    if stress_iters:
        assert rep["stress_iterations"] > 0
    # the global restarts are GPU calls of the anchor level over all anchors
    n_anc = [len(c["anchors"]) for c in case["chrs"]]
    assert sum(1 for c in calls if c["level"] == 2 and c["n"] in n_anc) >= 2 * params["restarts_anchor"]
    check_glue_calls(ctx, calls)


def test_ref_arcs_driver_band_16_seeds(ctx):
    case, rng = arcs_case(5)
    params = dict(PARAMS, restarts_anchor=2, restarts_subanchor=2, loop_density=2)
    pt0 = T.Tree(case["chrs"], case["splits"])
    ib_pos = np.cumsum(rng.normal(0, 1.0, size=(len(pt0.level(0, 2)), 3)), 0).astype(np.float32)
    ref_a, ref_s, prod_a, prod_s = [], [], [], []
    for seed in range(1, 17):
        print(f"arc driver band seed {seed}", flush=True)
        _, _, _, calls = ref_drive_arcs(case, params, seed, 0, ib_pos)
        if seed <= 2:
            check_glue_calls(ctx, calls)
        ref_a.append(block_scores(calls, 2))
        ref_s.append(block_scores(calls, 3))
        _, _, rep = product_drive_arcs(ctx, case, params, seed, 0, ib_pos)
        prod_a.append(rep["score_anchor"])
        prod_s.append(rep["score_smooth"])
    for name, r, q in (("anchor", ref_a, prod_a), ("smooth", ref_s, prod_s)):
        r, q = np.array(r), np.array(q)
        iqr = max(np.subtract(*np.percentile(r, [75, 25])), np.subtract(*np.percentile(q, [75, 25])))
        print(f"{name}: reference median {np.median(r):.4g} IQR {np.subtract(*np.percentile(r, [75, 25])):.3g}; "
              f"product median {np.median(q):.4g}", flush=True)
        assert abs(np.median(r) - np.median(q)) <= iqr, name
        assert sps.mannwhitneyu(r, q).pvalue > 1e-3, name


# ---------------------------------------------------------------- heatmap level
def heat_case():
    case, rng = genome(21, n_chr=2, n_anchors=(40, 30), span=3, splits_per_chr=4)
    pt = T.Tree(case["chrs"], case["splits"])
    return case, rng, pt


def heat_dist(pt, rng, level):
    n = sum(len(pt.level(c, level)) for c in range(2))
    idx = np.arange(n)
    f = (1.0 / (np.abs(idx[:, None] - idx[None, :]) ** 3 + 0.1)).astype(np.float32)
    f *= rng.uniform(0.5, 1.5, size=f.shape).astype(np.float32)
    f = ((f + f.T) / 2).astype(np.float32)
    dist, _ = T.distance_heatmap(f, 1, level == 0, product_params(PARAMS))
    return dist


def ref_drive_heatmap(case, params, level, dist, roots_pos, seed, segment):
    p = params
    text = header("drive_heatmap", p) + tree_text(case)
    text += (f"{level} {f32(p['noise_chr'])!r} {f32(p['noise_segment'])!r} {p['restarts_chr']} "
             f"{p['restarts_segment']} {p['use2d']} {FAILS}\n")
    text += sline(HEAT) + f"\n{seed} {segment} {CHAINS}\n"
    n = dist.shape[0]
    text += f"1 {n}\n" + " ".join(float(x).hex() for x in dist.reshape(-1)) + "\n"
    text += " ".join(float(x).hex() for x in np.asarray(roots_pos, np.float32).reshape(-1)) + "\n"
    lines, calls = run_ref(text)
    steps, active, cl = parse_drive(lines)
    return active, cl, calls


def product_drive_heatmap(ctx, case, params, level, dist, roots_pos, seed, segment):
    pt = T.Tree(case["chrs"], case["splits"])
    roots = [pt.chromosome(c)[1] for c in range(2)]
    pt.set_positions(roots, np.asarray(roots_pos, np.float32))
    po = T.heatmap_level_opts(heatmap=lvl_settings(HEAT), preset=1, chains_per_restart=CHAINS)
    rep = T.heatmap_level(ctx, pt, level, dist, 1, product_params(params), po, seed=seed,
                          segment=segment)
    return pt, rep


ROOTS = np.array([[0, 0, 0], [3, 1, 0]], np.float32)


@pytest.mark.parametrize("level", [0, 1])
def test_ref_heatmap_driver_noise_free_and_glue(ctx, level):
    case, rng, pt0 = heat_case()
    dist = heat_dist(pt0, rng, level)
    params = dict(PARAMS, noise_chr=0.0, noise_segment=0.0)
    print(f"reference heatmap driver level={level} n={dist.shape[0]}", flush=True)
    active, rcl, calls = ref_drive_heatmap(case, params, level, dist, ROOTS, 9, 40)
    pt, rep = product_drive_heatmap(ctx, case, params, level, dist, ROOTS, 9, 40)
    assert active == [c for ci in range(2) for c in pt.level(ci, level)]
    got = pt.positions(active)
    want = np.array([rcl[c]["pos"] for c in active], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert rep["score_anchor"] == pytest.approx(block_scores(calls, 1), rel=1e-6)
    check_glue_calls(ctx, calls)
    # with the noise on, the glue's calls still repeat exactly
    params = dict(PARAMS)
    _, _, calls = ref_drive_heatmap(case, params, level, dist, ROOTS, 9, 40)
    check_glue_calls(ctx, calls)


def test_ref_heatmap_driver_band_16_seeds(ctx):
    case, rng, pt0 = heat_case()
    dist = heat_dist(pt0, rng, 1)
    params = dict(PARAMS, restarts_segment=3)
    r, q = [], []
    for seed in range(1, 17):
        print(f"heatmap driver band seed {seed}", flush=True)
        _, _, calls = ref_drive_heatmap(case, params, 1, dist, ROOTS, seed, 0)
        r.append(block_scores(calls, 1))
        _, rep = product_drive_heatmap(ctx, case, params, 1, dist, ROOTS, seed, 0)
        q.append(rep["score_anchor"])
    r, q = np.array(r), np.array(q)
    iqr = max(np.subtract(*np.percentile(r, [75, 25])), np.subtract(*np.percentile(q, [75, 25])))
    print(f"heatmap: reference median {np.median(r):.4g}, product median {np.median(q):.4g}, "
          f"IQR {iqr:.3g}", flush=True)
    assert abs(np.median(r) - np.median(q)) <= iqr
    assert sps.mannwhitneyu(r, q).pvalue > 1e-3
