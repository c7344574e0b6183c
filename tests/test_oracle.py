"""Pins the FP64 oracle (oracle/) -- CPU only.

Each check is chosen so that a plausible mistake (dropped term, wrong sign or
index, transposed operand) fails it: published known-answer vectors, hand-
computed closed forms, invariants of the energies, identities between the
local and total forms, and the reference's own markArcs built from its
sources (oracle/_ref) or the fixtures it produced (tests/golden/).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import binding as O
from pmmc_amd.api import Arcs, Heatmap, Smooth

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------- Philox4x32-10 known-answer vectors ----------------
# Random123 kat_vectors (Salmon et al., SC'11), philox4x32 with 10 rounds.
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_kat(ctr, key, want):
    assert tuple(int(x) for x in O.philox(ctr, key)) == want


def test_sphere_penalty():
    L = O.load()
    p = np.array([3.0, 4.0, 0.0], np.float32)
    assert L.orc_sphere_penalty(p.ctypes.data_as(O.f32p), 0.0) == 0.0      # disabled
    assert L.orc_sphere_penalty(p.ctypes.data_as(O.f32p), 5.0) == 0.0      # on the surface
    assert L.orc_sphere_penalty(p.ctypes.data_as(O.f32p), 2.0) == pytest.approx(9.0)  # (5-2)^2


# ---------------- heatmap level ----------------
def _hm(pos, h, diag=1, bnd=None, thr=1e-6, fixed=None):
    return O.Problem(Heatmap(pos=np.asarray(pos, np.float32), dist=np.asarray(h, np.float32),
                             diagonal=diag, boundaries=bnd, skip_below=thr, fixed=fixed))


def test_heatmap_two_beads_closed_form():
    # h = 1, d = 2 -> each local term (2-1)/1 = 1 -> local 1, full 2
    P = _hm([[0, 0, 0], [2, 0, 0]], [[0, 1], [1, 0]])
    L = O.load()
    x = P.pos0
    assert L.orc_heatmap_local(O.C.byref(P.c), x.ctypes.data_as(O.f32p), 0) == pytest.approx(1.0)
    assert P.total() == pytest.approx(2.0)
    # d = 3, h = 2 -> (3-2)/2 = 0.5 -> 0.25 per term
    P = _hm([[0, 0, 0], [0, 3, 0]], [[0, 2], [2, 0]])
    assert P.total() == pytest.approx(0.5)


def test_heatmap_skip_rules():
    pos = [[0, 0, 0], [1, 0, 0], [3, 0, 0]]
    h = np.array([[0, 1, 2], [1, 0, 1], [2, 1, 0]], np.float32)
    # diag=1: all off-diagonal pairs: (1-1)^2, (3-2)^2/4, (2-1)^2 -> 0 + .25 + 1 = 1.25 per side
    assert _hm(pos, h, diag=1).total() == pytest.approx(2.5)
    # diag=2 drops |i-j| = 1: only (0,2): 0.25 per side
    assert _hm(pos, h, diag=2).total() == pytest.approx(0.5)
    # threshold drops h < thr
    h2 = h.copy()
    h2[0, 2] = h2[2, 0] = 1e-7
    assert _hm(pos, h2, diag=1).total() == pytest.approx(2.0)
    # chromosome blocks [0,1] and [2,2]: only pair (0,1)
    assert _hm(pos, h, diag=1, bnd=np.array([0, 2, 3], np.int32)).total() == pytest.approx(0.0)


def test_heatmap_invariants():
    rng = np.random.default_rng(3)
    n = 9
    pos = rng.normal(size=(n, 3)).astype(np.float32) * 3
    h = rng.uniform(0.5, 4, size=(n, n)).astype(np.float32)
    h = ((h + h.T) / 2).astype(np.float32)
    P = _hm(pos, h)
    e = P.total()
    # translation invariance
    assert _hm(pos + np.float32(5.5), h).total() == pytest.approx(e, rel=1e-5)
    # rotation invariance (90 degrees about z, exact in floats)
    rot = pos[:, [1, 0, 2]].copy()
    rot[:, 0] *= -1
    assert _hm(rot, h).total() == pytest.approx(e, rel=1e-9)
    # scale invariance: positions and targets scaled together (powers of two: exact)
    assert _hm(pos * 2, h * 2).total() == pytest.approx(e, rel=1e-9)
    # symmetric h: full = 2 * sum_{i<j}
    p64, h64 = pos.astype(np.float64), h.astype(np.float64)
    d = np.linalg.norm(p64[:, None] - p64[None], axis=2)
    iu = np.triu_indices(n, 1)
    assert e == pytest.approx(2 * np.sum(((d[iu] - h64[iu]) / h64[iu]) ** 2), rel=1e-9)


@pytest.mark.parametrize("symmetric", [True, False])
def test_heatmap_delta_is_full_difference(symmetric):
    rng = np.random.default_rng(11)
    n = 12
    pos = rng.normal(size=(n, 3)).astype(np.float32) * 2
    h = rng.uniform(0.5, 3, size=(n, n)).astype(np.float32)
    if symmetric:
        h = np.triu(h) + np.triu(h, 1).T
    h[rng.random((n, n)) < 0.1] = 0.0
    P = _hm(pos, h, diag=2, bnd=np.array([0, 5, 12], np.int32))
    for _ in range(20):
        p = int(rng.integers(n))
        q = pos[p] + rng.uniform(-1, 1, 3).astype(np.float32)
        y = pos.copy()
        y[p] = q
        assert P.delta(pos, p, q) == pytest.approx(P.total(y) - P.total(pos), rel=1e-9, abs=1e-9)


# ---------------- arc level ----------------
def _arcs(pos, ed, lp=None, lq=None, ks=1.0, kq=1.0):
    return O.Problem(Arcs(pos=np.asarray(pos, np.float32), exp_dist=np.asarray(ed, np.float32),
                          loop_pairs=lp, loop_params=lq, k_stretch=ks, k_squeeze=kq))


def test_arcs_closed_forms():
    ed = [[0, 1], [1, 0]]
    # stretched: d = 2, e = 1 -> ((2-1)/1)^2 * ks
    assert _arcs([[0, 0, 0], [2, 0, 0]], ed, ks=3.0, kq=7.0).total() == pytest.approx(3.0)
    # squeezed: d = 0.5 -> 0.25 * kq
    assert _arcs([[0, 0, 0], [0.5, 0, 0]], ed, ks=3.0, kq=7.0).total() == pytest.approx(1.75)
    # no arc: repulsion 1/d
    assert _arcs([[0, 0, 0], [4, 0, 0]], [[0, -1], [-1, 0]]).total() == pytest.approx(0.25)
    # explicit zero: ignored
    assert _arcs([[0, 0, 0], [4, 0, 0]], [[0, 0], [0, 0]]).total() == 0.0
    # loop k (d - r0)^2 on top of the repulsion
    P = _arcs([[0, 0, 0], [4, 0, 0]], [[0, -1], [-1, 0]], lp=np.array([[0, 1]]),
              lq=np.array([[2.0, 1.5]]))
    assert P.total() == pytest.approx(0.25 + 2.0 * 2.5 ** 2)


def test_arcs_local_matches_total_without_repulsion():
    rng = np.random.default_rng(5)
    n = 10
    pos = rng.normal(size=(n, 3)).astype(np.float32) * 2
    ed = np.full((n, n), -1.0, np.float32)
    np.fill_diagonal(ed, 0)
    for a, b in [(0, 3), (1, 2), (2, 7), (4, 9), (5, 6), (3, 8)]:
        ed[a, b] = ed[b, a] = rng.uniform(0.5, 3)
    ed[1, 5] = ed[5, 1] = 0.0
    lp = np.array([[0, 9], [2, 4]], np.int32)
    lq = np.array([[1.5, 2.0], [0.5, 1.0]], np.float32)
    P = _arcs(pos, ed, lp, lq, ks=1.3, kq=0.7)
    L = O.load()
    for _ in range(20):
        p = int(rng.integers(n))
        q = pos[p] + rng.uniform(-1, 1, 3).astype(np.float32)
        y = pos.copy()
        y[p] = q
        rep = (L.orc_arcs_repulsion(O.C.byref(P.c), y.ctypes.data_as(O.f32p))
               - L.orc_arcs_repulsion(O.C.byref(P.c), pos.ctypes.data_as(O.f32p)))
        assert P.delta(pos, p, q) == pytest.approx(P.total(y) - P.total(pos) - rep, rel=1e-9,
                                                   abs=1e-9)


# ---------------- smooth level ----------------
def _sm(pos, dtn=None, ks=0.1, kq=0.1, ka=0.1, wd=1.0, wa=1.0):
    pos = np.asarray(pos, np.float32)
    dtn = np.ones(len(pos), np.float32) if dtn is None else np.asarray(dtn, np.float32)
    return O.Problem(Smooth(pos=pos, dist_to_next=dtn, k_stretch=ks, k_squeeze=kq, k_angular=ka,
                            w_dist=wd, w_angle=wa))


def test_smooth_closed_forms():
    straight = [[i, 0, 0] for i in range(5)]
    assert _sm(straight).total() == 0.0                 # bonds at rest, no bending
    right = [[0, 0, 0], [1, 0, 0], [1, 1, 0]]
    assert _sm(right, ka=2.0).total() == pytest.approx(2.0 * 0.5 ** 3)   # ang = 1/2
    back = [[0, 0, 0], [1, 0, 0], [0, 0, 0]]
    assert _sm(back, ka=2.0).total() == pytest.approx(2.0)               # ang = 1
    stretched = [[0, 0, 0], [3, 0, 0]]
    assert _sm(stretched, ks=0.5).total() == pytest.approx(0.5 * 4.0)     # ((3-1)/1)^2
    # weights: distance term doubled, angle term tripled
    P = _sm([[0, 0, 0], [2, 0, 0], [2, 2, 0]], ks=1.0, ka=1.0, wd=2.0, wa=3.0)
    assert P.total() == pytest.approx(2.0 * (1.0 + 1.0) + 3.0 * 0.125)


def test_smooth_local_matches_unweighted_full():
    rng = np.random.default_rng(9)
    n = 12
    pos = np.cumsum(rng.normal(size=(n, 3)), axis=0).astype(np.float32)
    dtn = rng.uniform(0.5, 2.0, n).astype(np.float32)
    P = _sm(pos, dtn, ks=0.3, kq=0.2, ka=0.7, wd=1.7, wa=0.4)
    L = O.load()
    for p in range(n):
        q = pos[p] + rng.uniform(-1, 1, 3).astype(np.float32)
        y = pos.copy()
        y[p] = q
        full = lambda x: L.orc_smooth_full(O.C.byref(P.c), x.ctypes.data_as(O.f32p), 0.0, 0)
        assert P.delta(pos, p, q) == pytest.approx(full(y) - full(pos), rel=1e-9, abs=1e-12)


# ---------------- annealer schedule ----------------
def _sched(**kw):
    s = O.OrcSchedule()
    base = dict(T0=5.0, dt=0.999, improvement=0.99, small_energy=1e-6, ratio_stop=0.0,
                guard_energy=-math.inf, tjs=50.0, tjc=20.0, step=0.5, step_decay=1.0,
                sphere_radius=0.0, block=100, milestone_period=100, min_successes=10,
                fails_threshold=3, stop_rule=0, strict_accept=0, positive_guard=0,
                cool_per_block=0, use2d=0, resum=1, max_moves=1000)
    base.update(kw)
    for k, v in base.items():
        setattr(s, k, v)
    return s


def _small_heatmap():
    from pmmc_amd.synth import heatmap_problem
    return O.Problem(heatmap_problem(n=16, seed=2)[0])


def test_temperature_schedule_is_iterated_fp64():
    P = _small_heatmap()
    s = _sched(stop_rule=1, fails_threshold=10 ** 6, max_moves=700)
    _, st = P.anneal(s, seed=4, segment=0, chain=0)
    T = 5.0
    for _ in range(st["moves"]):
        T *= 0.999
    assert st["moves"] == 700 and st["T"] == T
    s = _sched(stop_rule=1, fails_threshold=10 ** 6, max_moves=700, cool_per_block=1)
    _, st = P.anneal(s, seed=4, segment=0, chain=0)
    assert st["T"] == 5.0 * 0.999 ** 7 or abs(st["T"] - 5.0 * 0.999 ** 7) < 1e-15


def test_cap_is_checked_before_each_block():
    P = _small_heatmap()
    s = _sched(stop_rule=1, fails_threshold=10 ** 6, block=256, milestone_period=256, max_moves=1000)
    _, st = P.anneal(s, seed=1, segment=0, chain=0)
    assert st["moves"] == 1024 and st["stop_reason"] == 3     # 4 blocks, like the GPU kernels


@pytest.mark.parametrize("k", [1, 3, 4])
def test_resum_interval(k):
    """resum = K replaces the running total by the full score at every K-th
    milestone only: after exactly K milestones E is the full score of the
    final positions; after K - 1 it is still the running total."""
    P = _small_heatmap()
    for blocks, expect_exact in ((k, True), (k - 1, False)):
        if blocks == 0:
            continue
        s = _sched(stop_rule=1, fails_threshold=10 ** 6, max_moves=100 * blocks, resum=k)
        x, st = P.anneal(s, seed=8, segment=0, chain=2)
        assert st["milestones"] == blocks
        run = _sched(stop_rule=1, fails_threshold=10 ** 6, max_moves=100 * blocks, resum=0)
        _, st0 = P.anneal(run, seed=8, segment=0, chain=2)
        if expect_exact:
            assert st["energy"] == P.total(x)
        else:
            assert st["energy"] == st0["energy"]


def test_zero_temperature_never_goes_uphill():
    P = _small_heatmap()
    s = _sched(T0=0.0, max_moves=3000, stop_rule=1, fails_threshold=10 ** 6, resum=0)
    x, st = P.anneal(s, seed=3, segment=0, chain=5, trace=True)
    assert st["energy"] <= P.total() + 1e-12
    assert st["accepted"] == int(st["trace"].sum())
    assert st["energy"] == pytest.approx(P.total(x), rel=1e-9)


def test_first_fail_stops_at_first_failing_milestone():
    P = _small_heatmap()
    # improvement = 0: E > 0*ms always; min_successes huge: always "no improvement"
    s = _sched(improvement=0.0, min_successes=10 ** 9, max_moves=10 ** 6)
    _, st = P.anneal(s, seed=3, segment=0, chain=0)
    assert st["moves"] == 100 and st["milestones"] == 1 and st["stop_reason"] == 1
    # MISSES_ACCUM with 3 fails: three milestones
    s = _sched(improvement=0.0, min_successes=10 ** 9, stop_rule=2, fails_threshold=3,
               max_moves=10 ** 6)
    _, st = P.anneal(s, seed=3, segment=0, chain=0)
    assert st["milestones"] == 3 and st["moves"] == 300


def test_guard_and_determinism():
    P = _small_heatmap()
    s = _sched(guard_energy=1e30)
    x, st = P.anneal(s, seed=3, segment=0, chain=0)
    assert st["stop_reason"] == 4 and st["moves"] == 0 and np.array_equal(x.reshape(-1), P.pos0)
    s = _sched()
    a, sa = P.anneal(s, seed=3, segment=0, chain=7)
    b, sb = P.anneal(s, seed=3, segment=0, chain=7)
    c, sc = P.anneal(s, seed=3, segment=0, chain=8)
    assert np.array_equal(a, b) and sa == sb and not np.array_equal(a, c)


def test_full_rescore_equals_incremental():
    P = _small_heatmap()
    s = _sched(resum=0, max_moves=500, stop_rule=1, fails_threshold=10 ** 6)
    a, sa = P.anneal(s, seed=2, segment=0, chain=1, trace=True)
    b, sb = P.anneal(s, seed=2, segment=0, chain=1, trace=True, full_rescore=1)
    assert np.array_equal(sa["trace"], sb["trace"]) and np.array_equal(a, b)
    assert sa["energy"] == pytest.approx(sb["energy"], rel=1e-9)


@pytest.mark.parametrize("level", ["heatmap", "arcs", "smooth"])
def test_near_tie_count(level):
    """The near-tie counter (tie_rel): off at 0; it never changes the chain;
    it grows with tie_rel; and at tie_rel = 1 every move counts, since |ΔE|
    never exceeds |local before| + |local after| (a mis-scaled bound fails
    this)."""
    from pmmc_amd import synth
    if level == "heatmap":
        P = _small_heatmap()
    elif level == "arcs":
        P = O.Problem(synth.arcs_problem(n=40, n_arcs=120, seed=3))
    else:
        P = O.Problem(synth.smooth_problem(n_anchors=8, add=4, seed=3))
    s = _sched(max_moves=1500, stop_rule=1, fails_threshold=10 ** 6, step=0.3)
    x0, s0 = P.anneal(s, seed=5, segment=0, chain=2, trace=True)
    assert s0["near_ties"] == 0 and s0["first_near_tie"] == -1
    counts = []
    for rel in (1e-7, 1e-5, 1e-3, 1.0):
        x, st = P.anneal(s, seed=5, segment=0, chain=2, trace=True, tie_rel=rel)
        assert np.array_equal(x, x0) and np.array_equal(st["trace"], s0["trace"])
        assert st["energy"] == s0["energy"] and st["moves"] == s0["moves"]
        counts.append(st["near_ties"])
        assert (st["near_ties"] == 0) == (st["first_near_tie"] == -1)
    assert counts == sorted(counts)
    assert counts[-1] == s0["moves"]


# ---------------- arc marking ----------------
def test_mark_arcs_closed_form():
    anc_s = [100, 300, 150, 0, 500, 700]
    anc_e = [200, 400, 350, 0, 500, 800]   # anchor 3 zero-length, anchor 4 length 1
    arc_s = [120, 160, 0, 500, 200, 700]
    arc_e = [310, 390, 100, 800, 300, 699]
    s, e = O.mark_arcs(anc_s, anc_e, arc_s, arc_e)
    assert s.tolist() == [0, 2, -1, -1, 2, 5]     # last containing anchor wins
    assert e.tolist() == [2, 1, 0, 5, 2, -1]      # 500 sits in a length-1 anchor: skipped


def _random_chromosome(rng, A=40, R=200, overlap=True):
    starts = np.sort(rng.integers(1, 100_000, size=A))
    lens = rng.integers(1, 3000 if overlap else 500, size=A)
    anc = np.stack([starts, starts + lens], 1)
    anc[rng.integers(0, A)] = (0, 0)
    anc = anc[np.argsort(anc[:, 0], kind="stable")]
    pick = rng.integers(0, A, size=(R, 2))
    gs = anc[pick[:, 0], 0] + rng.integers(0, 50, R)
    ge = anc[pick[:, 1], 0] + rng.integers(0, 50, R)
    score = rng.integers(1, 20, R)
    factor = rng.integers(0, 3, R)
    raw = np.stack([gs, ge, score, factor], 1)
    raw = np.concatenate([raw, raw[:R // 5]])   # duplicates -> merged groups
    return anc.astype(np.int32), raw.astype(np.int32)


def test_group_map_order_matches_reference_markArcs():
    """The std::map variant against the reference's own CPU markArcs."""
    rng = np.random.default_rng(21)
    chrs = [_random_chromosome(rng) for _ in range(3)]
    ref = O.ref_mark_arcs(chrs)
    if ref is None:
        fx = json.load(open(os.path.join(GOLDEN, "markarcs_ref.json")))
        chrs = [(np.array(c["anchors"], np.int32), np.array(c["raw"], np.int32)) for c in fx["chromosomes"]]
        ref = [np.array(c["arcs"], np.int32).reshape(-1, 7) for c in fx["chromosomes"]]
    ours, _ = O.group_arcs(chrs, use_unordered=False)
    for a, b in zip(ours, ref):
        assert np.array_equal(a, b)


def test_group_unordered_is_same_multiset():
    rng = np.random.default_rng(22)
    chrs = [_random_chromosome(rng) for _ in range(2)]
    u, nu = O.group_arcs(chrs, use_unordered=True)
    m, nm = O.group_arcs(chrs, use_unordered=False)
    assert np.array_equal(nu, nm)
    for a, b in zip(u, m):
        ka = sorted(map(tuple, a.tolist()))
        kb = sorted(map(tuple, b.tolist()))
        assert ka == kb


def test_markarcs_golden_fixture_matches_oracle():
    fx = json.load(open(os.path.join(GOLDEN, "markarcs_ref.json")))
    chrs = [(np.array(c["anchors"], np.int32), np.array(c["raw"], np.int32)) for c in fx["chromosomes"]]
    ours, _ = O.group_arcs(chrs, use_unordered=False)
    for a, c in zip(ours, fx["chromosomes"]):
        assert np.array_equal(a, np.array(c["arcs"], np.int32).reshape(-1, 7))
