"""CPU restatement of the host-side rows around the annealing kernels --
TEST INFRASTRUCTURE ONLY (imported by tests/, never by pmmc_amd/).

Restates, in plain Python with float32 arithmetic where the reference computes
in float (numpy float32 scalars round every operation to float32, as C does;
powf/pow/exp are libm's, through ctypes):
  build_tree        LooperSolver::createTreeGenome / createTreeChromosome
                    (LooperSolver.hpp:2272-2438), findGaps (:2119-2180),
                    findSplit (:2185-2269)
  set_level         LooperSolver::setTopLevel / setLevel / levelDown
                    (:2431-2452, :3087-3104)
  densify           LooperSolver::densifyActiveRegion (:3836-3898)
  dist_to_next      dist_to_next / noise_size of reconstructClusterArcsDistances
                    (:4339-4354) with genomicLengthToDistance (:3900-3904)
  freq_to_distance  LooperSolver::freqToDistance (:3949-3967)
  expected_distances LooperSolver::calcAnchorExpectedDistancesHeatmap
                    (:5494-5627)
  distance_heatmap  LooperSolver::createDistanceHeatmap (:3042-3085) with
                    freqToDistanceHeatmap(Inter) (:3934-3942) and
                    Heatmap::getDiagonalSize / getAvg / getRange
                    (Heatmap.hpp:356-367, :427-451)
  stress_input      the set-up of LooperSolver::stressRefineAnchors
                    (:5642-5760): sorted active region, pair list, spring
                    targets (the descent itself is oracle/stress.c)
  reinterpolate     the sub-anchor re-placement of the global anchor
                    refinement (:4179-4239)

PINNED BY tests/golden/levels_ref.json, written by the reference's own code
(oracle/_ref/levels_ref), and by the invariants in tests/test_levels.py
(children partition parents, genomic monotonicity, base ranges, counts).
The driver loops arcs_level / heatmap_level: the product's drivers match them
bit for bit (tests/test_gpu_levels.py) and match the reference's own drivers,
run through the reference-side glue, bit for bit at zero restart noise and
within a 16-seed band with noise (tests/test_gpu_ref_drivers.py).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

_libm = ctypes.CDLL("libm.so.6")
_libm.powf.restype = ctypes.c_float
_libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
F = np.float32
INT_MAX = 2**31 - 1


class ReferenceError_(RuntimeError):
    """The reference calls error() (prints and exits) for this input."""


def c_div(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def new_cluster(start, end):
    """Cluster::Cluster(int, int) (Cluster.hpp:115-136)."""
    return dict(start=start, end=end, genomic_pos=c_div(start + end, 2), orientation="N",
                base_start=0, base_end=0, parent=-1, is_fixed=False, level=0,
                children=[], arcs=[], pos=[F(0), F(0), F(0)], dist_to_next=0.0)


# ------------------------------------------------------------------ T1
def _find_gaps(t, chr_name):
    """LooperSolver::findGaps (LooperSolver.hpp:2119-2180)."""
    cl, arcs = t["clusters"], t["arcs"][chr_name]
    start = t["chr_first"][chr_name]
    gaps = [start]
    cnt = 0
    for i in range(start, len(cl)):
        for a in cl[i]["arcs"]:
            s, e = arcs[a][0], arcs[a][1]
            other = e if s == i else (s if e == i else -1)
            if other == i:
                continue
            cnt += 1 if other > i else -1
        if cnt == 0:
            gaps.append(i)
    if len(cl) - 1 not in gaps:
        gaps.append(len(cl) - 1)
    regions = t["splits"]
    if regions:
        for rc, _, rend in regions:
            if rc != chr_name:
                continue
            for i in range(start, len(cl)):
                if cl[i]["end"] >= rend:
                    if i not in gaps:
                        gaps.append(i)
                    break
        gaps = sorted(set(gaps))
    return gaps


def _find_split(t, gaps, chr_name):
    """LooperSolver::findSplit (LooperSolver.hpp:2185-2269)."""
    cl, regions = t["clusters"], t["splits"]
    if len(cl) < len(gaps):
        raise ReferenceError_(f"find split: size mismatch, gaps={len(gaps)}, clusters={len(cl)}")
    if not regions:
        return gaps
    splits = [gaps[0]]
    first, last = -1, len(regions) - 1
    for i, (rc, _, _) in enumerate(regions):
        if rc == chr_name:
            if first == -1:
                first = i
        elif first != -1:
            last = i - 1
            break
    if first == -1:
        raise ReferenceError_(f"no segment split found for {chr_name}")
    for s in range(first, last):
        boundary = regions[s][2]
        best, best_d = -1, INT_MAX
        for i in range(1, len(gaps) - 1):
            d = abs(cl[gaps[i]]["end"] - boundary)
            if d < best_d:
                best, best_d = i, d
        if best >= 0 and gaps[best] not in splits:
            splits.append(gaps[best])
    splits.sort()
    if gaps[-1] not in splits:
        splits.append(gaps[-1])
    return splits


def build_tree(chrs, splits):
    """createTreeGenome (LooperSolver.hpp:2272-2293) over createTreeChromosome
    (:2296-2438).  chrs: [{name, anchors [(start, end)], orient, arcs
    [(a, b, score)] with chromosome-local anchor indices}]; splits: predefined
    segments [(chr, start, end)] in file order."""
    t = dict(clusters=[], arcs={}, scores={}, chr_first={}, chr_root={}, chrs=[c["name"] for c in chrs],
             splits=[tuple(s) for s in splits], anchors={},
             long_arcs={c["name"]: [tuple(x) for x in c.get("long_arcs", [])] for c in chrs})
    cl = t["clusters"]
    for c in chrs:
        name = c["name"]
        cs = len(cl)
        t["chr_first"][name] = cs
        t["anchors"][name] = [tuple(a) for a in c["anchors"]]
        for i, (s, e) in enumerate(c["anchors"]):
            k = new_cluster(s, e)
            k["orientation"] = c.get("orient", ["N"] * len(c["anchors"]))[i]
            k["base_start"] = k["base_end"] = cs + i
            k["level"] = 4
            cl.append(k)
        t["arcs"][name] = [[a + cs, b + cs] for a, b, _ in c["arcs"]]
        t["scores"][name] = [sc for _, _, sc in c["arcs"]]
        for i, (a, b) in enumerate(t["arcs"][name]):
            cl[a]["arcs"].append(i)
            cl[b]["arcs"].append(i)
        root = new_cluster(0, 0)
        gaps = _find_gaps(t, name)
        spl = _find_split(t, gaps, name)
        next_split = 1
        start_gap = len(cl)
        for i in range(1, len(gaps)):
            prev_gap = gaps[i - 1] if i == 1 else gaps[i - 1] + 1
            cur_gap = gaps[i]
            k = new_cluster(cl[prev_gap]["start"], cl[cur_gap]["end"])
            k["base_start"], k["base_end"], k["level"] = prev_gap, cur_gap, 3
            for j in range(prev_gap, cur_gap + 1):
                k["children"].append(j)
                cl[j]["parent"] = len(cl)
            cl.append(k)
            if gaps[i] == spl[next_split]:
                end_gap = len(cl) - 1
                s = new_cluster(cl[start_gap]["start"], cl[end_gap]["end"])
                s["base_start"] = cl[start_gap]["base_start"]
                s["base_end"] = cl[end_gap]["base_end"]
                s["level"] = 2
                seg = len(cl)
                for j in range(start_gap, end_gap + 1):
                    s["children"].append(j)
                    cl[j]["parent"] = seg
                root["children"].append(seg)
                cl.append(s)
                start_gap = len(cl)
                next_split += 1
        if not root["children"]:
            raise ReferenceError_(f"chromosome {name}: no segment (reads children[0] of an empty list)")
        root["level"] = 1
        root["start"] = cl[root["children"][0]]["start"]
        root["end"] = cl[root["children"][-1]]["end"]
        root["genomic_pos"] = c_div(root["start"] + root["end"], 2)
        root["base_start"] = cl[root["children"][0]]["base_start"]
        root["base_end"] = cl[root["children"][-1]]["base_end"]
        for ch in root["children"]:
            cl[ch]["parent"] = len(cl)
        t["chr_root"][name] = len(cl)
        cl.append(root)
    return t


def set_level(t, level):
    """setLevel = setTopLevel + level x levelDown (LooperSolver.hpp:2431-2452,
    :3087-3104); children keep their stored order."""
    cur = {c: [t["chr_root"][c]] for c in t["chrs"]}
    for _ in range(level):
        for c in t["chrs"]:
            nxt = []
            for k in cur[c]:
                ch = t["clusters"][k]["children"]
                nxt.extend(ch if ch else [k])
            cur[c] = nxt
    return cur


# ------------------------------------------------------------------ S5
def densify(t, ib, loop_density, fix):
    """densifyActiveRegion(ib, fix) (LooperSolver.hpp:3836-3898) with
    active_region = children of ib.  Appends the new beads to t's clusters and
    returns the new active region (also stored as ib's children)."""
    cl = t["clusters"]
    active = list(cl[ib]["children"])
    if not active:
        return active
    add = loop_density
    level = cl[active[0]]["level"] + 1
    out = []
    for i in range(len(active) - 1):
        a, b = cl[active[i]], cl[active[i + 1]]
        if fix:
            a["is_fixed"] = True
        rng = b["start"] - a["end"]
        d = c_div(rng, add + 1)
        p = a["end"]
        out.append(active[i])
        dst = F(F(1.0) / F(add + 1))
        st = dst
        for _ in range(add):
            p += d
            k = new_cluster(p, p)
            k["pos"] = [F(pa + F(F(pb - pa) * st)) for pa, pb in zip(a["pos"], b["pos"])]
            k["level"] = level
            cl.append(k)
            out.append(len(cl) - 1)
            st = F(st + dst)
    out.append(active[-1])
    if fix:
        cl[active[-1]]["is_fixed"] = True
    cl[ib]["children"] = list(out)
    return out


def genomic_length_to_distance(length, p):
    """genomicLengthToDistance (LooperSolver.hpp:3900-3904): double math with
    float settings."""
    return float(F(p["glen_base"])) + float(F(p["glen_scale"])) * math.pow(length / 1000.0,
                                                                        float(F(p["glen_power"])))


def dist_to_next(genomic_pos, p, coef_subanchor=0.5, coef_anchor=0.5):
    """dist_to_next and noise_size of reconstructClusterArcsDistances
    (LooperSolver.hpp:4339-4354).  Returns (float32 dists, noise for the
    subanchor coefficient, noise for the anchor coefficient)."""
    out = []
    noise = F(0.0)
    for i in range(len(genomic_pos) - 1):
        d = abs(genomic_pos[i + 1] - genomic_pos[i])
        dist = F(genomic_length_to_distance(d, p))
        out.append(dist)
        noise = F(noise + dist)
    noise = F(noise / F(len(genomic_pos)))
    return np.array(out, np.float32), F(noise * F(coef_subanchor)), F(noise * F(coef_anchor))


# ------------------------------------------------------------------ A6
def freq_to_distance(freq, p):
    """LooperSolver::freqToDistance (LooperSolver.hpp:3949-3967); the memo of
    counts 2..100 holds the same values (:950-951)."""
    a, scale, shift, base = F(p["count_a"]), F(p["count_scale"]), F(p["count_shift"]), F(p["count_base"])
    if a < 0.0:
        fi = int(F(F(freq) + shift))        # float -> int truncates toward zero
        fi = max(fi, 1)
        return F(float(base) + float(scale) * math.pow(float(fi), float(a)))
    x = float(F(F(freq) + shift))
    try:
        e = math.exp(float(a) * x)
    except OverflowError:      # C's exp returns +inf
        e = math.inf
    return F(float(base) + float(scale) / e)


def diagonal_size(h):
    """Heatmap::getDiagonalSize (Heatmap.hpp:356-367)."""
    n = h.shape[0]
    for w in range(n):
        for i in range(n - w):
            if float(h[i, i + w]) > 1e-6:
                return w
    return 0


def heat_range(h):
    """Heatmap::getRange (Heatmap.hpp:427-438) -> (min, max)."""
    mx, mn = F(0.0), F(1e9)
    w = diagonal_size(h)
    n = h.shape[0]
    for i in range(n):
        for j in range(i + w, n):
            v = h[i, j]
            if v < mn:
                mn = v
            if v > mx:
                mx = v
    return mn, mx


def heat_avg(h):
    """Heatmap::getAvg (Heatmap.hpp:440-451): sequential float sum."""
    n = h.shape[0]
    w = diagonal_size(h)
    s = F(0.0)
    cnt = 0
    for i in range(n):
        for j in range(i + w, n):
            s = F(s + h[i, j])
            cnt += 1
    with np.errstate(invalid="ignore", divide="ignore"):
        return F(s / F(cnt))


def expected_distances(t, chr_name, ib, long_arcs, p, anchor_heatmap=None, influence=0.5,
                       active=None):
    """calcAnchorExpectedDistancesHeatmap (LooperSolver.hpp:5494-5627) with
    active_region = children of ib (or `active`) and current_chr = chr_name.
    An arc whose other end lies outside the active region lands in slot 0, as
    the reference's std::map::operator[] default does (:5538-5539)."""
    if long_arcs is None:        # the chromosome's long arcs stored in the tree
        long_arcs = t["long_arcs"].get(chr_name, [])
    cl = t["clusters"]
    if active is None:
        active = cl[ib]["children"]
    n = len(active)
    m = np.full((n, n), -1.0, np.float32)
    for i in range(n):
        m[i, i] = 0.0
    c2a = {}
    for i, c in enumerate(active):
        c2a[c] = i
    arcs, scores = t["arcs"][chr_name], t["scores"][chr_name]
    for i, ai in enumerate(active):
        for a in cl[ai]["arcs"]:
            s, e = arcs[a]
            other = e if s == ai else (s if e == ai else -1)
            if other == -1:
                raise ReferenceError_("no other end")
            if other < ai:
                continue
            ed = freq_to_distance(scores[a], p)
            x = c2a[ai]
            y = c2a.setdefault(other, 0)
            m[x, y] = ed
            m[y, x] = ed
    anc = t["anchors"].get(chr_name)
    if long_arcs and anc is not None:
        n_anc = len(anc)
        cs = t["chr_first"][chr_name]
        a2a = [-1] * n_anc
        for ci in sorted(c2a):
            if cs <= ci < cs + n_anc:
                a2a[ci - cs] = c2a[ci]

        def pos_to_anc(x):
            for k, (s, e) in enumerate(anc):
                if s <= x <= e:
                    return k
            return -1

        for s, e, sc in long_arcs:
            ia, ib2 = pos_to_anc(s), pos_to_anc(e)
            if ia < 0 or ib2 < 0 or ia == ib2:
                continue
            x, y = a2a[ia], a2a[ib2]
            if x < 0 or y < 0:
                continue
            if m[x, y] >= 0.0:
                continue
            ed = freq_to_distance(sc, p)
            m[x, y] = ed
            m[y, x] = ed
    if anchor_heatmap is not None:
        h = np.asarray(anchor_heatmap, np.float32)
        if h.shape != (n, n):
            raise ReferenceError_("anchor heatmaps size mismatch!")
        _, mx = heat_range(h)
        if mx > 1e-6:
            infl = F(influence)
            for i in range(n):
                for j in range(i + 1, n):
                    if m[i, j] < 0.0:
                        continue
                    s = F(F(h[i, j] / mx) * infl)
                    if s > 1.0:
                        s = F(1.0)
                    m[i, j] = F(m[i, j] * F(F(1.0) - s))
                    m[j, i] = m[i, j]
    return m


# ------------------------------------------------------------------ H6
def distance_heatmap(freq, diag, inter, p):
    """createDistanceHeatmap (LooperSolver.hpp:3042-3085); inter selects
    freqToDistanceHeatmapInter (heatmap index 0).  Returns (h, avg used for
    the clamp)."""
    f = np.asarray(freq, np.float32)
    n = f.shape[0]
    scale = F(p["heat_scale_inter"] if inter else p["heat_scale"])
    power = F(p["heat_power_inter"] if inter else p["heat_power"])
    h = np.zeros((n, n), np.float32)
    for i in range(n):
        for j in range(i, n):
            v = f[i, j]
            if float(v) < 1e-6:
                x = F(0.0)
            elif abs(i - j) < diag:
                x = F(-1.0)
            else:
                x = F(scale * F(_libm.powf(float(v), float(power))))
            h[i, j] = x
            h[j, i] = x
    avg = heat_avg(h)
    mx = F(avg * F(p["stretching"]))
    h[h > mx] = mx
    return h, avg


# ------------------------------------------------------------------ H7 / A7 drivers
LEVEL_INIT = 5


def restart_noise(seed, segment, bead, amp, use2d):
    """random_vector(amp, use2d) (common.cpp:14-25) on the Philox stream the
    library uses for restart noise: (2u-1)*amp per axis from
    Philox(seed, {bead, 0, segment, 5<<28})."""
    from oracle import binding as O
    o = O.philox([bead, 0, segment, LEVEL_INIT << 28], [seed & 0xFFFFFFFF, seed >> 32])
    out = [F(F(F(2.0) * F(F(int(x) >> 8) * F(2.0 ** -24)) - F(1.0)) * F(amp)) for x in o[:3]]
    if use2d:
        out[2] = F(0.0)
    return out


def _offer(best, b, score, payload, gpu_member=True):
    """best-of rule score < best || best < 0 (LooperSolver.hpp:1392-1398,
    :4457-4463) on the value the annealing member returns: a float from the
    GPU members, a double from the CPU annealer"""
    if gpu_member:
        score = float(np.float32(score))
    cur = best.get(b)
    if cur is None or score < cur[0] or cur[0] < 0.0:
        best[b] = (score, payload)


def call_segments(sizes, restarts_anchor, restarts_subanchor, loop_density, segment,
                  chroms=None, global_anchors=None):
    """Philox segments of the anchor and sub-anchor restarts of each block, in
    the order the reference driver issues its GPU calls (LooperSolver.hpp:
    4027-4138; blocks of at most one child make none; sub-anchor problems
    below 4 beads run the CPU annealer, :4431-4437, and take segments from
    2^30 on).  sizes: children per block before densification.  With
    global_anchors ({chromosome: anchors}, in chromosome order) and chroms
    (each block's chromosome), the global refinement's restarts follow each
    chromosome's blocks (:4141-4170) and ({chromosome: first segment}) is
    returned as well."""
    call, cpu, out = 0, 0, []
    gseg = {}

    def block(n):
        nonlocal call, cpu
        if n <= 1:
            out.append((0, 0))
            return
        a = segment + call
        call += restarts_anchor
        if n + (n - 1) * loop_density >= 4:
            sm = segment + call
            call += restarts_subanchor
        else:
            sm = segment + (1 << 30) + cpu
            cpu += restarts_subanchor
        out.append((a, sm))

    if global_anchors is None:
        for n in sizes:
            block(n)
        return out
    k = 0
    for name, n_anc in global_anchors.items():
        while k < len(sizes) and chroms[k] == name:
            block(sizes[k])
            k += 1
        if n_anc > 1:
            gseg[name] = segment + call
            call += restarts_anchor
    return out, gseg


def global_stage(t, chr_list, p, o, seed, gseg, limit, run_job, rep):
    """The global anchor refinement of reconstructClustersArcsDistances'
    chromosome loop (LooperSolver.hpp:4141-4243) for each listed chromosome:
    restarts of the arc annealer over all its anchors from their current
    positions (restart noise noise_anchor_small on every bead, step = mean
    dist_to_next x noise_anchor), the best restart kept; then, when
    o["stress_iters"] > 0, stressRefineAnchors (stress_input + the oracle's
    descent); then the sub-anchor re-placement among the clusters that existed
    when the reference ran it (limit[chromosome])."""
    from oracle import binding as O
    cl = t["clusters"]
    todo = [name for name in chr_list if len(t["anchors"][name]) > 1]
    best = {}
    for name in todo:
        cs = t["chr_first"][name]
        active = list(range(cs, cs + len(t["anchors"][name])))
        ed = expected_distances(t, name, None, None, p, active=active)
        dtn, _, noise = dist_to_next([cl[c]["genomic_pos"] for c in active], p,
                                     p["noise_subanchor"], p["noise_anchor"])
        for i, c in enumerate(active[:-1]):
            cl[c]["dist_to_next"] = float(dtn[i])
        sched = o["arcs_sched"](noise)
        for k in range(p["restarts_anchor"]):
            seg = gseg[name] + k
            pos = np.array([[F(cl[c]["pos"][a] + restart_noise(seed, seg, i, p["noise_anchor_small"],
                                                               p["use2d"])[a]) for a in range(3)]
                            for i, c in enumerate(active)], np.float32)
            prob = dict(pos=pos, fixed=np.array([cl[c]["is_fixed"] for c in active], np.uint8),
                        exp_dist=ed, k_stretch=o["k_stretch_arcs"], k_squeeze=o["k_squeeze_arcs"])
            r = run_job(2, prob, sched, seed, seg, o["chains"])
            rep["n_jobs"] += 1
            _offer(best, name, r["energy"], (active, r))
    rep.setdefault("score_global", 0.0)
    for name, (score, (active, r)) in best.items():
        for i, c in enumerate(active):
            cl[c]["pos"] = [F(x) for x in r["pos"][i]]
        rep["score_global"] += score
    rep.setdefault("stress_iterations", 0)
    rep.setdefault("reinterpolated", 0)
    orders = {}
    for name in todo:
        cs = t["chr_first"][name]
        act = list(range(cs, cs + len(t["anchors"][name])))
        if o.get("stress_iters", 0) > 0:
            act, p0, bb, pairs = stress_input(t, name, act, p, o["k_stretch"], o["k_squeeze"])
            out, st = O.stress_refine(p0, bb, pairs, o["k_stretch"], o["k_squeeze"], o["k_angular"],
                                      o["stress_iters"], o["stress_lr"])
            for i, c in enumerate(act):
                cl[c]["pos"] = [F(x) for x in out[i]]
            rep["stress_iterations"] += st["iterations"]
        orders[name] = act
    for name in todo:
        rep["reinterpolated"] += reinterpolate(t, orders[name], limit[name])


def arcs_level(t, chr_name, p, o, seed, segment, run_job):
    """reconstructClustersArcsDistances' block loop (LooperSolver.hpp:3978-4140)
    over reconstructClusterArcsDistances (:4303-4487), in the library's
    batched order (job segments numbered in the reference's call order):
    every block's anchor restarts, then densification, then
    every block's sub-anchor restarts.  p: level parameters (dict, the keys of
    pmmc_level_params), o: dict with arcs_sched(step) / smooth_sched(step,
    n_beads) schedule factories, springs, chains and densify.  run_job(level,
    problem dict, schedule, seed, segment, n_chains) returns
    dict(pos, energy, resummed) of the job's best chain.  Returns the report
    dict and records candidates per block in t["_candidates"]."""
    cl = t["clusters"]
    blocks = []
    for name in t["chrs"]:
        if chr_name is not None and name != chr_name:
            continue
        for ib in set_level(t, 2)[name]:
            if cl[ib]["level"] == 3:
                blocks.append((name, ib))
    rep = dict(n_blocks=len(blocks), n_jobs=0, score_anchor=0.0, score_smooth=0.0)
    chr_list = [name for name in t["chrs"] if chr_name is None or name == chr_name]
    glob = bool(o.get("global_refine", False))
    sizes = [len(cl[ib]["children"]) for _, ib in blocks]
    if glob:
        segs, gseg = call_segments(sizes, p["restarts_anchor"], p["restarts_subanchor"],
                                   p["loop_density"], segment, chroms=[n for n, _ in blocks],
                                   global_anchors={n: len(t["anchors"][n]) for n in chr_list})
    else:
        segs = call_segments(sizes, p["restarts_anchor"], p["restarts_subanchor"],
                             p["loop_density"], segment)
    limit = {name: len(cl) for name in chr_list}
    best = {}
    for bi, (name, ib) in enumerate(blocks):
        active = list(cl[ib]["children"])
        for c in active:
            cl[c]["pos"] = list(cl[ib]["pos"])
        if len(active) <= 1:
            continue
        ed = expected_distances(t, name, ib, None, p)
        dtn, _, noise = dist_to_next([cl[c]["genomic_pos"] for c in active], p,
                                     p["noise_subanchor"], p["noise_anchor"])
        for i, c in enumerate(active[:-1]):
            cl[c]["dist_to_next"] = float(dtn[i])
        sched = o["arcs_sched"](noise)
        for k in range(p["restarts_anchor"]):
            seg = segs[bi][0] + k
            pos = np.array([[F(cl[c]["pos"][a] + restart_noise(seed, seg, i, p["noise_anchor_small"],
                                                               p["use2d"])[a]) for a in range(3)]
                            for i, c in enumerate(active)], np.float32)
            prob = dict(pos=pos, fixed=np.array([cl[c]["is_fixed"] for c in active], np.uint8),
                        exp_dist=ed, k_stretch=o["k_stretch_arcs"], k_squeeze=o["k_squeeze_arcs"])
            r = run_job(2, prob, sched, seed, seg, o["chains"])
            rep["n_jobs"] += 1
            _offer(best, bi, r["energy"], (active, r))
    cand = {}
    for bi, (score, (active, r)) in best.items():
        for i, c in enumerate(active):
            cl[c]["pos"] = [F(x) for x in r["pos"][i]]
        rep["score_anchor"] += score
    if not o.get("densify", True):
        if glob:
            global_stage(t, chr_list, p, o, seed, gseg, limit, run_job, rep)
        return rep
    for name in chr_list:
        for bname, ib in blocks:
            if bname == name:
                densify(t, ib, p["loop_density"], True)
        limit[name] = len(cl)
    best = {}
    for bi, (name, ib) in enumerate(blocks):
        active = list(cl[ib]["children"])
        if len(active) <= 1:
            continue
        dtn, noise, _ = dist_to_next([cl[c]["genomic_pos"] for c in active], p,
                                     p["noise_subanchor"], p["noise_anchor"])
        for i, c in enumerate(active[:-1]):
            cl[c]["dist_to_next"] = float(dtn[i])
        sched = o["smooth_sched"](noise, len(active))
        for k in range(p["restarts_subanchor"]):
            seg = segs[bi][1] + k
            pos = []
            for i, c in enumerate(active):
                d = [F(0), F(0), F(0)] if cl[c]["is_fixed"] else restart_noise(seed, seg, i, noise, p["use2d"])
                pos.append([F(cl[c]["pos"][a] + d[a]) for a in range(3)])
            prob = dict(pos=np.array(pos, np.float32),
                        fixed=np.array([cl[c]["is_fixed"] for c in active], np.uint8),
                        dist_to_next=np.array([cl[c]["dist_to_next"] for c in active], np.float32),
                        k_stretch=o["k_stretch"], k_squeeze=o["k_squeeze"], k_angular=o["k_angular"],
                        w_dist=o["w_dist"], w_angle=o["w_angle"])
            r = run_job(3, prob, sched, seed, seg, o["chains"])
            rep["n_jobs"] += 1
            cand.setdefault(bi, []).append((r["resummed"], active, r))
            _offer(best, bi, r["resummed"], (active, r), len(active) >= 4)
    for bi, (score, (active, r)) in best.items():
        for i, c in enumerate(active):
            if not cl[c]["is_fixed"]:
                cl[c]["pos"] = [F(x) for x in r["pos"][i]]
        rep["score_smooth"] += score
    t["_candidates"] = cand
    if glob:
        global_stage(t, chr_list, p, o, seed, gseg, limit, run_job, rep)
    return rep


def heatmap_level(t, level, dist, diag, p, o, seed, segment, run_job):
    """reconstructClustersHeatmapSingleLevel (LooperSolver.hpp:1258-1422):
    restarts of the heatmap annealer over the concatenated level lists, with
    the library's restart noise; o: dict with sched(step), chains,
    skip_below."""
    cl = t["clusters"]
    cur = set_level(t, level)
    active = [c for name in t["chrs"] for c in cur[name]]
    n = len(active)
    avg = heat_avg(np.asarray(dist, np.float32))
    avg_dist = F(avg * F(p["noise_chr"] if level == 0 else p["noise_segment"]))
    bnd = None
    if level == 1:
        for name in t["chrs"]:
            par = cl[cur[name][0]]["parent"]
            for s in cur[name]:
                cl[s]["pos"] = list(cl[par]["pos"])
        if len(t["chrs"]) > 1:
            bnd = [0]
            for name in t["chrs"]:
                bnd.append(bnd[-1] + len(cur[name]))
    restarts = p["restarts_chr"] if level == 0 else p["restarts_segment"]
    sched = o["sched"](avg_dist)
    best = {}
    results = []
    for k in range(restarts):
        seg = segment + k
        pos = np.array([[F(cl[c]["pos"][a] + restart_noise(seed, seg, i, avg_dist, p["use2d"])[a])
                         for a in range(3)] for i, c in enumerate(active)], np.float32)
        prob = dict(pos=pos, fixed=np.array([cl[c]["is_fixed"] for c in active], np.uint8),
                    dist=np.asarray(dist, np.float32), diagonal=diag,
                    boundaries=None if bnd is None else np.array(bnd, np.int32),
                    skip_below=o["skip_below"])
        r = run_job(1, prob, sched, seed, seg, o["chains"])
        results.append((r["resummed"], r))
        _offer(best, 0, r["resummed"], r)
    score, r = best[0]
    for i, c in enumerate(active):
        cl[c]["pos"] = [F(x) for x in r["pos"][i]]
    t["_candidates"] = {0: [(s, active, rr) for s, rr in results]}
    return dict(n_blocks=1, n_jobs=restarts, score_anchor=score, avg_dist=avg_dist)


# ---------------------------------------------------------------------------
# heatmap input construction (upstream of createDistanceHeatmap)
# ---------------------------------------------------------------------------
MAX_POSITION = 2000000000  # LooperSolver.hpp:2554
LAST_BREAK = 1000000000    # :2536


def singleton_breaks(clusters):
    """Bin boundaries of one chromosome (LooperSolver.hpp:2527-2536): 0, the
    midpoints between consecutive clusters of the level, 1e9."""
    return [0] + [c_div(clusters[i][1] + clusters[i + 1][0], 2)
                  for i in range(len(clusters) - 1)] + [LAST_BREAK]


def _first_bin(breaks, pos):
    """Linear scan of the reference (:2652-2657): the first i with
    breaks[i] <= pos <= breaks[i+1], or -1."""
    for i in range(len(breaks) - 1):
        if breaks[i] <= pos <= breaks[i + 1]:
            return i
    return -1


def _first_bins(breaks, pos):
    """_first_bin for an array of positions (numpy, same first-match rule)."""
    b = np.asarray(breaks, np.int64)
    pos = np.asarray(pos, np.int64)
    out = np.full(pos.shape, -1, np.int64)
    for i in range(len(b) - 2, -1, -1):   # the lowest matching i is written last
        hit = (b[i] <= pos) & (pos <= b[i + 1])
        out[hit] = i
    return out


def length_excluding_centromere(cent, st, end):
    """LooperSolver::getGenomicLengthExcludingCentromere (:5916-5935), cent =
    the chromosome's first centromere region (start, end) or None."""
    if cent is None:
        return end - st + 1
    cs, ce = cent
    if cs > end or ce < st:
        return end - st + 1
    if cs < st:
        if ce < end:
            return end - ce + 1
        return 0
    if ce < end:
        return end - st - (ce - cs + 1) + 1
    return cs - st


def singleton_heatmap(chrs, clusters, records, region=(0, 0), centromeres=(), long_arcs=(),
                      long_scale=10.0, long_power=2.0, diag=1):
    """LooperSolver::createSingletonHeatmap (:2517-2749) on parsed records.
    chrs: chromosome names in order; clusters: per chromosome the level's
    (start, end) in level order; records: (chr1, sta, stb, chr2, enda, endb,
    score); centromeres: (chr, start, end) regions (the first per chromosome
    counts); long_arcs: (chr, start, end, score).  Returns (heat [N,N]
    float32, counts dict).  Float sums in record order, as the reference."""
    breaks, start_ind, cur = {}, {}, 0
    for c, cl in zip(chrs, clusters):
        breaks[c] = singleton_breaks(cl)
        start_ind[c] = cur
        cur += len(breaks[c]) - 1
    n = cur
    h = np.zeros((n, n), np.float32)
    cmax = {c: 0 for c in chrs}
    cmin = {c: MAX_POSITION for c in chrs}
    sel = region[1] > 0
    cnt = dict(added=0, diagonal=0, bad_chr=0)
    rec = list(records)
    idx = {c: k for k, c in enumerate(chrs)}
    c1 = np.array([idx.get(r[0], -1) for r in rec], np.int64)
    c2 = np.array([idx.get(r[3], -1) for r in rec], np.int64)
    sta, stb, enda, endb, sc = (np.array([r[k] for r in rec], np.int64) for k in (1, 2, 4, 5, 6))
    good = (c1 >= 0) & (c2 >= 0)
    cnt["bad_chr"] = int((~good).sum())
    if sel:
        good &= (sta >= region[0]) & (sta <= region[1]) & (endb >= region[0]) & (endb <= region[1])
    sta = np.where(sta < 0, 0, sta)
    enda = np.where(enda < 0, 0, enda)
    st = np.full(len(rec), -1, np.int64)
    en = np.full(len(rec), -1, np.int64)
    for c, k in idx.items():
        m = good & (c1 == k)
        st[m] = _first_bins(breaks[c], sta[m]) + np.where(_first_bins(breaks[c], sta[m]) >= 0, start_ind[c], 0)
        m = good & (c2 == k)
        en[m] = _first_bins(breaks[c], enda[m]) + np.where(_first_bins(breaks[c], enda[m]) >= 0, start_ind[c], 0)
    if np.any(good & ((st < 0) | (en < 0))):
        raise ReferenceError_("non matching arc (the reference exits)")
    dg = good & (st == en)
    cnt["diagonal"] = int(dg.sum())
    ok = good & (st != en)
    cnt["added"] = int(ok.sum())
    for c, k in idx.items():
        for mc, lo, hi in ((c1 == k, sta, stb), (c2 == k, enda, endb)):
            m = ok & mc
            if m.any():
                cmax[c] = max(cmax[c], int(hi[m].max()))
                cmin[c] = min(cmin[c], int(lo[m].min()))
    # float accumulation in record order (np.add.at is unbuffered and ordered)
    np.add.at(h, (st[ok], en[ok]), sc[ok].astype(np.float32))
    np.add.at(h, (en[ok], st[ok]), sc[ok].astype(np.float32))
    for c in chrs:
        for (ac, a, b, s) in long_arcs:
            if ac != c:
                continue
            i, j = _first_bin(breaks[c], a), _first_bin(breaks[c], b)
            if i < 0 or j < 0:
                raise ReferenceError_("long arc outside the bins (the reference indexes -1)")
            i += start_ind[c]
            j += start_ind[c]
            if i == j:
                continue
            val = F(float(F(long_scale)) * math.pow(float(s), float(F(long_power))))
            h[i, j] = F(h[i, j] + val)   # twice into [st][end] (:2735-2736)
            h[i, j] = F(h[i, j] + val)
    cent = {}
    for (c, s, e) in centromeres:
        cent.setdefault(c, (s, e))
    lens = []
    for c in chrs:
        b = breaks[c]
        b[0] = cmin[c]
        b[-1] = cmax[c]
        for i in range(len(b) - 1):
            d = length_excluding_centromere(cent.get(c) if centromeres else None, b[i], b[i + 1])
            lens.append(F(F(d) / F(1e6)))
    lens = np.array(lens, np.float32)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        iu = np.triu_indices(n, 1)
        v = h[iu] / (lens[iu[0]] * lens[iu[1]])
    out = np.zeros_like(h)
    out[iu] = v
    out[(iu[1], iu[0])] = v
    for i in range(n):
        for p in range(diag):
            if i + p < n:
                out[i, i + p] = 0.0
                out[i + p, i] = 0.0
    return out, cnt


def normalize_heatmap(heat):
    """LooperSolver::normalizeHeatmap (:2996-3040): every row scaled to the
    average row sum (rows summing to <= 1e-12 become 0), then symmetrised as
    the mean of the two triangles.  Sequential float sums."""
    v = np.asarray(heat, np.float32)
    n = v.shape[0]
    rows = np.array([np.add.accumulate(v[i], dtype=np.float32)[-1] if n else F(0) for i in range(n)],
                    np.float32)
    expected = np.add.accumulate(rows, dtype=np.float32)[-1] if n else F(0)
    expected = F(expected / F(n))
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        mn = np.where(rows > F(1e-12), expected / rows, F(0)).astype(np.float32)
        h = (v * mn[:, None]).astype(np.float32)
    iu = np.triu_indices(n, 1)
    s = ((h[iu] + h[(iu[1], iu[0])]) / F(2)).astype(np.float32)
    h[iu] = s
    h[(iu[1], iu[0])] = s
    return h


def normalize_heatmap_diagonal_total(heat, val=1.0):
    """LooperSolver::normalizeHeatmapDiagonalTotal (:3146-3172): scale so that
    the first non-zero diagonal averages `val` (FP64 average); unchanged when
    that average is not above 1e-12."""
    h = np.array(heat, np.float32)
    n = h.shape[0]
    w = diagonal_size(h)
    avg = 0.0
    for i in range(n - w):
        avg += float(h[i, i + w])
    avg /= (n - w)
    if not avg > 1e-12:
        return h
    mn = F(float(F(val)) / avg)
    return (h * mn).astype(np.float32)


def normalize_heatmap_inter(heat, sizes, scale):
    """LooperSolver::normalizeHeatmapInter (:3106-3144): every cell times
    `scale`, then the intra-chromosome blocks divided by it again (float);
    unchanged for one chromosome."""
    h = np.array(heat, np.float32)
    if len(sizes) <= 1:
        return h
    s = F(scale)
    h = (h * s).astype(np.float32)
    o = 0
    for k in sizes:
        h[o:o + k, o:o + k] = (h[o:o + k, o:o + k] / s).astype(np.float32)
        o += k
    return h


# ------------------------------------------------------------------ stress refinement
def stress_input(t, chr_name, active, p, ks, kq):
    """stressRefineAnchors' set-up (LooperSolver.hpp:5642-5760): the active
    region sorted by genomic position (a stable sort: the reference's
    std::sort agrees where positions differ), pairs from the chromosome's arcs
    then its long arcs (first pair of two anchors kept; none between
    neighbours in the sorted order while a spring constant exceeds 1e-6; none
    with a target <= 1e-6), spring targets from the genomic gaps (-1 for a
    zero gap).  Returns (sorted active, positions [N,3], spring targets,
    pairs [(i, j, t)])."""
    cl = t["clusters"]
    active = sorted(active, key=lambda c: cl[c]["genomic_pos"])
    N = len(active)
    anc = t["anchors"][chr_name]
    n_anc, cs = len(anc), t["chr_first"][chr_name]
    a2a = [-1] * n_anc
    for i, c in enumerate(active):
        if cs <= c < cs + n_anc:
            a2a[c - cs] = i
    backbone = F(ks) > F(1e-6) or F(kq) > F(1e-6)
    pairs, seen = [], set()

    def add(a, b, tgt):
        if a == b:
            return
        a, b = min(a, b), max(a, b)
        if backbone and b - a <= 1:
            return
        if (a, b) in seen:
            return
        seen.add((a, b))
        pairs.append((a, b, tgt))

    for (a, b), sc in zip(t["arcs"][chr_name], t["scores"][chr_name]):
        if not (cs <= a < cs + n_anc and cs <= b < cs + n_anc):
            continue
        ai, aj = a2a[a - cs], a2a[b - cs]
        if ai < 0 or aj < 0:
            continue
        tt = freq_to_distance(sc, p)
        if tt > F(1e-6):
            add(ai, aj, tt)

    def pos_to_anc(x):
        for k, (s0, e0) in enumerate(anc):
            if s0 <= x <= e0:
                return k
        return -1

    for s0, e0, sc in t["long_arcs"].get(chr_name, []):
        ia, ib = pos_to_anc(s0), pos_to_anc(e0)
        if ia < 0 or ib < 0:
            continue
        a, b = a2a[ia], a2a[ib]
        if a < 0 or b < 0:
            continue
        tt = freq_to_distance(sc, p)
        if tt > F(1e-6):
            add(a, b, tt)
    bb = np.full(max(N - 1, 1), -1.0, np.float32)
    for i in range(N - 1):
        gap = abs(cl[active[i + 1]]["genomic_pos"] - cl[active[i]]["genomic_pos"])
        if gap > 0:
            bb[i] = F(genomic_length_to_distance(gap, p))
    pos = np.array([cl[c]["pos"] for c in active], np.float32).reshape(-1, 3)
    return active, pos, bb, pairs


def reinterpolate(t, active, limit):
    """The global refinement's sub-anchor re-placement (LooperSolver.hpp:
    4179-4239): every bead of level (anchors' level + 1) among the first
    `limit` clusters whose genomic position lies within [first, last] of
    `active` (its order) moves onto the segment between the anchors around
    it.  Returns the number of beads moved."""
    import bisect
    cl = t["clusters"]
    if not active:
        return 0
    sub = cl[active[0]]["level"] + 1
    gp = [cl[c]["genomic_pos"] for c in active]
    moved = 0
    for ci in range(min(limit, len(cl))):
        b = cl[ci]
        if b["level"] != sub or b["genomic_pos"] < gp[0] or b["genomic_pos"] > gp[-1]:
            continue
        g = b["genomic_pos"]
        j = bisect.bisect_left(gp, g)
        if j == 0:
            b["pos"] = [F(x) for x in cl[active[0]]["pos"]]
        elif j >= len(gp):
            b["pos"] = [F(x) for x in cl[active[-1]]["pos"]]
        else:
            g0, g1 = gp[j - 1], gp[j]
            u = F(F(g - g0) / F(g1 - g0)) if g1 > g0 else F(0.5)
            p0, p1 = cl[active[j - 1]]["pos"], cl[active[j]]["pos"]
            b["pos"] = [F(F(F(F(1.0) - u) * F(p0[a])) + F(u * F(p1[a]))) for a in range(3)]
        moved += 1
    return moved
