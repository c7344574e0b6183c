"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Per-move ΔE tolerance (stated per level): |ΔE_gpu - ΔE_fp64| <=
1e-5 * (|local_before| + |local_after|) + 1e-7, i.e. 1e-5 relative to the
magnitude of the two local sums the delta is formed from (the delta itself
is a difference and can be arbitrarily small).  Trajectories: the kernels and
the oracle consume the same Philox stream, so chain positions are compared
bit-for-bit; a chain may only differ when an FP32 vs FP64 accept decision
flips on a near-tie, which the test bounds.
"""
import numpy as np
import pytest

from oracle import binding as O
from pmmc_amd import synth
from pmmc_amd.api import Arcs, Heatmap, Smooth, schedule_preset, settings

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _proposals(rng, n, k, scale):
    beads = rng.integers(0, n, size=k).astype(np.int32)
    disp = rng.uniform(-scale, scale, size=(k, 3)).astype(np.float32)
    return beads, disp


def _check_delta(ctx, prob, k=2048, scale=0.5, r=0.0, seed=0, label=None):
    rng = np.random.default_rng(seed)
    P = O.Problem(prob)
    n = P.n
    beads, disp = _proposals(rng, n, k, scale)
    got = ctx.delta_energy(prob, beads, disp, sphere_radius=r)
    pos = P.pos0.reshape(-1, 3)
    worst = 0.0
    for i in range(k):
        p = int(beads[i])
        q = pos[p] + disp[i]
        want = P.delta(pos, p, q, r)
        a, b = P.local_pair(pos, p, q, r)
        err = abs(got[i] - want) / (abs(a) + abs(b) + 1e-7 / TOL)
        worst = max(worst, err)
    print(f"ΔE worst relative error [{label or type(prob).__name__} n={n} r={r}]: {worst:.3e}",
          flush=True)
    assert worst <= TOL, f"worst relative ΔE error {worst:.3e}"
    return worst


# The measured worst ΔE error per level, at each level's workload size, on
# proposals of the level's own step scale (stated per level; the bar is TOL).
@pytest.mark.parametrize("level", [1, 2, 3])
def test_delta_error_per_level_at_workload_size(ctx, level):
    prob, step, _ = synth.config_case(level, 0)
    scale = {1: 0.75, 2: 1.0, 3: 1.0}[level] * step
    worst = _check_delta(ctx, prob, k=4096, scale=scale, seed=level,
                         label=f"level {level} workload")
    assert worst <= TOL


def test_delta_heatmap(ctx):
    prob, _ = synth.heatmap_problem(n=64, seed=1)
    _check_delta(ctx, prob)
    prob, _ = synth.heatmap_problem(n=150, seed=2, diagonal=2)
    prob.boundaries = np.array([0, 60, 150], np.int32)
    _check_delta(ctx, prob, r=4.0)


def test_delta_heatmap_asymmetric(ctx):
    rng = np.random.default_rng(3)
    n = 40
    h = rng.uniform(0.5, 5, (n, n)).astype(np.float32)
    np.fill_diagonal(h, 0)
    prob = Heatmap(pos=rng.normal(size=(n, 3)).astype(np.float32) * 3, dist=h)
    _check_delta(ctx, prob)


def test_delta_arcs(ctx):
    prob = synth.arcs_problem(n=400, n_arcs=1200, seed=1)
    prob.loop_pairs = np.array([[0, 10], [5, 300], [7, 7]], np.int32)
    prob.loop_params = np.array([[1.0, 1.5], [0.5, 3.0], [1.0, 1.0]], np.float32)
    prob.k_stretch, prob.k_squeeze = 1.3, 0.8
    _check_delta(ctx, prob, r=6.0)


def test_delta_smooth(ctx):
    prob = synth.smooth_problem(n_anchors=40, add=9, seed=1)
    prob.w_dist, prob.w_angle = 1.5, 0.5
    _check_delta(ctx, prob, scale=0.3, r=20.0)


def test_delta_smooth_nearly_straight(ctx):
    n = 64
    pos = np.stack([np.arange(n, dtype=np.float32), np.zeros(n, np.float32),
                    np.zeros(n, np.float32)], 1)
    pos += np.random.default_rng(4).normal(scale=1e-3, size=pos.shape).astype(np.float32)
    prob = Smooth(pos=pos, dist_to_next=np.ones(n, np.float32), k_angular=5.0)
    _check_delta(ctx, prob, scale=0.01)


# ---------------- trajectories ----------------
def _run_pair(ctx, prob, sched, seed=7, segment=3, n_chains=64):
    """GPU plain anneal (64 chains = one ensemble) vs the oracle's 64 chains."""
    pos, st = ctx.anneal(prob, sched, seed=seed, segment=segment, chain0=0, n_chains=n_chains)
    P = O.Problem(prob)
    osched = O.schedule_from(sched)
    res = [P.anneal(osched, seed, segment, c) for c in range(n_chains)]
    energies = np.array([r[1]["energy"] for r in res])
    best = int(np.argmin(energies))
    return pos, st, res, best


def _assert_traj(pos, st, res, best, total_moves_rel=0.0):
    ox, ost = res[st["best_chain"]]
    assert st["best_chain"] == best, (st["best_chain"], best)
    assert np.array_equal(pos, ox), "best chain positions differ from the oracle's chain"
    assert st["best_chain_moves"] == ost["moves"]
    assert st["best_chain_accepted"] == ost["accepted"]
    assert st["energy"] == pytest.approx(ost["energy"], rel=1e-6)
    assert st["moves"] == sum(r[1]["moves"] for r in res)


# 33 <= n <= 64 runs the register-resident kernels (symmetric weight tables:
# the lane-pair form; asymmetric: one lane per chain); n <= 32 and n = 100
# the partner loop over memory.
@pytest.mark.parametrize("n,asym,radius,use2d,preset", [
    (32, False, 0.0, 0, 0), (32, False, 0.0, 0, 1), (64, False, 0.0, 0, 1), (50, True, 0.0, 0, 1),
    (40, False, 6.0, 1, 1), (20, True, 5.0, 0, 0), (100, False, 0.0, 0, 1)])
def test_trajectory_heatmap(ctx, n, asym, radius, use2d, preset):
    prob, _ = synth.heatmap_problem(n=n, seed=5)
    if asym:   # scale the upper triangle: h[i][j] != h[j][i] where both are targets
        rng = np.random.default_rng(n)
        up = np.triu(prob.dist > 0, 1)
        prob.dist = np.where(up, prob.dist * rng.uniform(0.8, 1.2, prob.dist.shape),
                             prob.dist).astype(np.float32)
    s = settings(max_temp=5.0, dt_temp=0.999, improvement=0.99, steps=512, min_successes=10,
                 milestone_fails=3, sphere_radius=radius, use2d=use2d)
    sched = schedule_preset(1, preset, s, 1.0)
    sched.max_moves = 4096
    print(f"trajectory heatmap n={n} asym={asym} r={radius} 2d={use2d} preset={preset}", flush=True)
    _assert_traj(*_run_pair(ctx, prob, sched))


@pytest.mark.parametrize("preset", [0, 1])
def test_trajectory_arcs(ctx, preset):
    prob = synth.arcs_problem(n=120, n_arcs=360, seed=5)
    s = settings(max_temp=5.0, dt_temp=0.9999, improvement=0.975, steps=2048, min_successes=100,
                 milestone_fails=3)
    sched = schedule_preset(2, preset, s, 0.3)
    sched.max_moves = 4096
    _assert_traj(*_run_pair(ctx, prob, sched))


@pytest.mark.parametrize("preset", [0, 1])
def test_trajectory_smooth(ctx, preset):
    prob = synth.smooth_problem(n_anchors=20, add=9, seed=5)
    s = settings(max_temp=5.0, dt_temp=0.9995, improvement=0.99, steps=2048, min_successes=50,
                 milestone_fails=3)
    sched = schedule_preset(3, preset, s, 0.5)
    sched.max_moves = 4096
    pos, st, res, best = _run_pair(ctx, prob, sched)
    ox, ost = res[st["best_chain"]]
    assert st["best_chain"] == best
    free = prob.fixed == 0
    assert np.array_equal(pos[free], ox[free])
    assert np.array_equal(pos[~free], prob.pos[~free])
    assert st["energy"] == pytest.approx(ost["energy"], rel=1e-6)


def test_resummed_energy_matches_oracle(ctx):
    prob, _ = synth.heatmap_problem(n=48, seed=9)
    s = settings(steps=512)
    sched = schedule_preset(1, 1, s, 1.0)
    sched.max_moves = 2048
    pos, st = ctx.anneal(prob, sched, seed=1, n_chains=64)
    P = O.Problem(prob)
    # heatmap return value: the full score of the returned positions
    assert st["energy_resummed"] == pytest.approx(P.total(pos), rel=1e-12)


# ---------------- arc marking ----------------
def _chrom(rng, A, R, sort=True):
    starts = rng.integers(1, 300_000, size=A)
    if sort:
        starts = np.sort(starts)
    lens = rng.integers(0, 5000, size=A)
    anc = np.stack([starts, starts + lens], 1)
    anc[rng.integers(0, A, size=3)] = 0
    pick = rng.integers(0, A, size=(R, 2))
    gs = np.abs(anc[pick[:, 0], 0] + rng.integers(-20, 80, R))
    ge = np.abs(anc[pick[:, 1], 0] + rng.integers(-20, 80, R))
    raw = np.stack([gs, ge, rng.integers(1, 30, R), rng.integers(0, 3, R)], 1)
    return anc.astype(np.int32), raw.astype(np.int32)


@pytest.mark.parametrize("sort", [True, False])
def test_mark_arcs_bit_exact(ctx, sort):
    rng = np.random.default_rng(31 + sort)
    anc, raw = _chrom(rng, 3000, 20000, sort)
    s, e = ctx.mark_arcs(anc[:, 0], anc[:, 1], raw[:, 0], raw[:, 1])
    os_, oe = O.mark_arcs(anc[:, 0], anc[:, 1], raw[:, 0], raw[:, 1])
    assert np.array_equal(s, os_) and np.array_equal(e, oe)


def test_mark_arcs_edges(ctx):
    s, e = ctx.mark_arcs([], [], [1, 2], [3, 4])
    assert s.tolist() == [-1, -1] and e.tolist() == [-1, -1]
    s, e = ctx.mark_arcs([5], [9], [], [])
    assert s.size == 0


def test_group_arcs_bit_exact(ctx):
    rng = np.random.default_rng(41)
    chrs = [_chrom(rng, 200, 1500) for _ in range(4)]
    got, nm = ctx.group_arcs(chrs)
    want, wnm = O.group_arcs(chrs, use_unordered=True)
    assert np.array_equal(nm, wnm)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_group_arcs_golden_fixture(ctx):
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "markarcs_ref.json")))
    chrs = [(np.array(c["anchors"], np.int32), np.array(c["raw"], np.int32)) for c in fx["chromosomes"]]
    got, _ = ctx.group_arcs(chrs)
    for a, c in zip(got, fx["chromosomes"]):
        ref = np.array(c["arcs"], np.int32).reshape(-1, 7)
        assert sorted(map(tuple, a.tolist())) == sorted(map(tuple, ref.tolist()))


# ---------------- batched == plain (alone vs in the mix) ----------------
def test_batch_equals_plain(ctx):
    s = settings(max_temp=5.0, dt_temp=0.9999, improvement=0.975, steps=2048, min_successes=100)
    sched = schedule_preset(2, 1, s, 0.3)
    sched.max_moves = 3072
    probs = [synth.arcs_problem(n=n, n_arcs=3 * n, seed=50 + n) for n in (30, 90, 200)]
    hs, _ = synth.heatmap_problem(n=40, seed=3)
    hsched = schedule_preset(1, 1, settings(steps=512), 1.0)
    hsched.max_moves = 2048
    alone = [ctx.anneal(p, sched, seed=11, segment=i, n_chains=128) for i, p in enumerate(probs)]
    alone.append(ctx.anneal(hs, hsched, seed=11, segment=3, n_chains=64))
    mixed = ctx.anneal_batch(probs + [hs], [sched] * 3 + [hsched], seeds=[11] * 4,
                             n_chains=[128, 128, 128, 64], blocks_per_step=3, use_graph=True)
    for (pa, sa), (pb, sb) in zip(alone, mixed):
        assert np.array_equal(pa, pb)
        assert sa["energy"] == sb["energy"] and sa["moves"] == sb["moves"]
        assert sa["best_chain"] == sb["best_chain"]


@pytest.mark.parametrize("cpw", [64, 32])
def test_batch_one_launch_equals_plain(ctx, cpw):
    """blocks_per_step = -1: one launch per level group runs every chain to
    its end (no steps, graph or tuner); same results as plain calls, with one
    lane or a lane pair per arc chain."""
    s = settings(max_temp=5.0, dt_temp=0.9999, improvement=0.975, steps=2048, min_successes=100)
    sched = schedule_preset(2, 1, s, 0.3)
    sched.max_moves = 3072
    probs = [synth.arcs_problem(n=n, n_arcs=3 * n, seed=60 + n) for n in (25, 110, 300)]
    ss = schedule_preset(3, 1, settings(steps=1024, dt_temp=0.9995), 0.5)
    sm = synth.smooth_problem(n_anchors=12, add=9, seed=5)
    # a 48-bead heatmap runs the queued register kernel, a 100-bead one the
    # partner loop over memory
    hs = schedule_preset(1, 1, settings(steps=512), 1.0)
    hs.max_moves = 2048
    h48, _ = synth.heatmap_problem(n=48, seed=6)
    h100, _ = synth.heatmap_problem(n=100, seed=7)
    alone = [ctx.anneal(p, sched, seed=13, segment=i, n_chains=192) for i, p in enumerate(probs)]
    alone.append(ctx.anneal(sm, ss, seed=13, segment=3, n_chains=64))
    alone.append(ctx.anneal(h48, hs, seed=13, segment=4, n_chains=128))
    alone.append(ctx.anneal(h100, hs, seed=13, segment=5, n_chains=64))
    shape = {}
    one = ctx.anneal_batch(probs + [sm, h48, h100], [sched] * 3 + [ss, hs, hs], seeds=[13] * 6,
                           n_chains=[192, 192, 192, 64, 128, 64], blocks_per_step=-1,
                           chains_per_wave=cpw, opts_out=shape)
    assert shape["blocks_per_step"] == -1 and shape["chains_per_wave"] == cpw
    assert shape["tuned"] == 0
    for (pa, sa), (pb, sb) in zip(alone, one):
        assert np.array_equal(pa, pb)
        assert sa["energy"] == sb["energy"] and sa["moves"] == sb["moves"]


# ---------------- heatmap sweep (reference GPU heatmap algorithm) ----------------
# The GPU scores a proposal in FP32 (reciprocal weights, hardware square
# root, two interleaved accumulators) and keeps each bead's reference score
# as a float; the oracle scores in FP64.  A decision can differ only when the
# oracle finds it within TIE_SWEEP (relative) of its threshold: the
# proposal's downhill or Metropolis test, the argmin over proposers, or (in
# FP64 on both sides) the candidate-vs-snapshot and stop tests.
TIE_SWEEP = 2e-6


def _sweep_scheds(iters, inner=4, proposers=64):
    from pmmc_amd import _abi as A
    args = (5.0, 0.9999, 0.99, 1e-4, 50.0, 20.0, 0.75, 0.95, 0.0, inner, proposers, 3, iters, 0)
    return A.SweepSchedule(*args), O.OrcSweep(*args)


@pytest.mark.parametrize("n,seed", [(24, 9), (24, 3), (40, 5)])
def test_sweep_matches_oracle(ctx, n, seed):
    from pmmc_amd import _abi as A
    prob, _ = synth.heatmap_problem(n=n, seed=6, skip_below=1e-3)
    P = O.Problem(prob)
    checked = 0
    for k in range(1, 6):
        ss, osw = _sweep_scheds(k)
        print(f"sweep n={n} seed={seed} iterations={k}", flush=True)
        pos, st = ctx.heatmap_sweep(prob, ss, seed=seed, segment=1)
        x, ost = O.sweep(P, osw, seed, 1, tie_rel=TIE_SWEEP)
        first = ost["first_near_tie"]
        if first < 0:
            # no decision of the k iterations was near its threshold: the
            # runs are the same, bit for bit
            assert np.array_equal(pos, x), k
            assert st["milestones"] == ost["milestones"]
            assert st["stop_reason"] == A.STOP_REASONS[ost["stop_reason"]]
            assert st["energy"] == pytest.approx(ost["energy"], rel=1e-12)
            checked += n
        elif first == k - 1:
            # the first near tie is in the last iteration: only the beads it
            # touched may differ, every other bead is bit-exact
            diff = np.any(pos != x, axis=1)
            assert not np.any(diff & ~ost["tie_beads"]), (k, np.nonzero(diff & ~ost["tie_beads"]))
            assert ost["tie_beads"].sum() <= n // 4
            checked += int((~ost["tie_beads"]).sum())
        else:
            # an earlier near tie may have changed the snapshot of later
            # iterations: the runs agree statistically only
            assert st["energy"] == pytest.approx(ost["energy"], rel=2e-2)
        assert st["energy"] < P.total()          # the sweep lowers the full score
    assert checked >= n - n // 4      # most beads compared bit for bit


# ---------------- served form ----------------
# (blocks_per_step, launch_ms, streams): fixed blocks per step; adaptive
# pacing on one stream with short launches; adaptive on 16 streams
SERVER_SHAPES = [(2, 0.0, 0), (0, 0.5, 1), (0, 30.0, 16)]


@pytest.mark.parametrize("bps,launch_ms,streams", SERVER_SHAPES)
def test_server_alone_equals_mixed(ctx, bps, launch_ms, streams):
    s = settings(max_temp=5.0, dt_temp=0.9999, improvement=0.975, steps=2048, min_successes=100)
    asched = schedule_preset(2, 1, s, 0.3)
    asched.max_moves = 2048
    hsched = schedule_preset(1, 1, settings(steps=512), 1.0)
    hsched.max_moves = 1024
    ssched = schedule_preset(3, 1, settings(steps=1024, dt_temp=0.9995), 0.5)
    jobs = [(synth.arcs_problem(n=n, n_arcs=3 * n, seed=70 + n), asched, 128) for n in (40, 150)]
    jobs.append((synth.heatmap_problem(n=30, seed=4)[0], hsched, 64))
    jobs.append((synth.smooth_problem(n_anchors=12, add=9, seed=2), ssched, 64))
    alone = [ctx.anneal(p, sc, seed=5, segment=i, n_chains=nc) for i, (p, sc, nc) in enumerate(jobs)]
    srv = ctx.server(hbm_budget=1 << 30, max_jobs=3, blocks_per_step=bps, launch_ms=launch_ms,
                     streams=streams)
    ids = [srv.submit(p, sc, seed=5, segment=i, n_chains=nc) for i, (p, sc, nc) in enumerate(jobs)]
    max_res, steps = 0, 0
    while not all(srv.done(j) for j in ids):
        _, r, _ = srv.step()
        max_res = max(max_res, r)
        steps += 1
        assert steps < 10000
    assert max_res <= 3
    for (pa, sa), j in zip(alone, ids):
        pb, sb = srv.result(j)
        assert np.array_equal(pa, pb)
        assert sa["energy"] == sb["energy"] and sa["best_chain"] == sb["best_chain"]
    srv.close()


def test_server_submit_many_equals_plain(ctx):
    """Jobs of one pmmc_server_submit_many call that name the same problem
    share its preparation; each result equals the plain call's, including a
    job list that interleaves two problems with their own schedules, chain
    counts and seeds."""
    hs = schedule_preset(1, 1, settings(steps=512), 1.0)
    hs.max_moves = 1024
    asched = schedule_preset(2, 1, settings(steps=2048, improvement=0.975, min_successes=100), 0.3)
    asched.max_moves = 2048
    hp = synth.heatmap_problem(n=48, seed=9)[0]
    ap = synth.arcs_problem(n=120, n_arcs=360, seed=71)
    srv = ctx.server(hbm_budget=1 << 30, max_jobs=64, blocks_per_step=0)
    ids_h = srv.submit_many(hp, hs, seeds=[3, 4, 5], segment=7, n_chains=128)
    ids_a = srv.submit_many(ap, asched, seeds=[6, 7], segment=2, n_chains=64)
    while not all(srv.done(j) for j in ids_h + ids_a):
        srv.step()
    for sd, j in zip([3, 4, 5], ids_h):
        pa, sa = ctx.anneal(hp, hs, seed=sd, segment=7, n_chains=128)
        pb, sb = srv.result(j)
        assert np.array_equal(pa, pb) and sa["energy"] == sb["energy"], sd
    for sd, j in zip([6, 7], ids_a):
        pa, sa = ctx.anneal(ap, asched, seed=sd, segment=2, n_chains=64)
        pb, sb = srv.result(j)
        assert np.array_equal(pa, pb) and sa["energy"] == sb["energy"], sd
    # a bad job in the list queues nothing
    bad = schedule_preset(2, 1, settings(steps=2048), 0.3)
    bad.block = 0
    with pytest.raises(Exception):
        srv.submit_many(ap, bad, seeds=[1, 2])
    _, r, q = srv.step()
    assert r == 0 and q == 0
    srv.close()


def test_batch_shared_problem_equals_distinct(ctx):
    """A batch whose jobs name one problem object (prepared once) equals the
    same batch over separate copies of it (prepared per job)."""
    import copy
    asched = schedule_preset(2, 1, settings(steps=2048, improvement=0.975, min_successes=100), 0.3)
    asched.max_moves = 2048
    ap = synth.arcs_problem(n=90, n_arcs=270, seed=72)
    hp = synth.heatmap_problem(n=40, seed=3)[0]
    hs = schedule_preset(1, 1, settings(steps=512), 1.0)
    hs.max_moves = 1024
    shared = [ap, hp, ap, ap, hp]
    distinct = [copy.deepcopy(p) for p in shared]
    scheds = [asched, hs, asched, asched, hs]
    seeds = [1, 2, 3, 4, 5]
    a = ctx.anneal_batch(shared, scheds, seeds=seeds, n_chains=64)
    b = ctx.anneal_batch(distinct, scheds, seeds=seeds, n_chains=64)
    for (pa, sa), (pb, sb) in zip(a, b):
        assert np.array_equal(pa, pb) and sa["energy"] == sb["energy"]


def test_batch_shared_problem_errors_name_the_job(ctx):
    """Jobs sharing a problem's preparation still fail one by one, as
    separately prepared jobs would: a bad schedule on the first job (the one
    whose tables the others copy) or on a later one is reported for that job,
    and a bad problem for the first job naming it."""
    from pmmc_amd.api import PmmcError
    good = schedule_preset(2, 1, settings(steps=1024), 0.3)
    good.max_moves = 1024
    bad = schedule_preset(2, 1, settings(steps=1024), 0.3)
    bad.block = 0
    ap = synth.arcs_problem(n=50, n_arcs=150, seed=73)
    with pytest.raises(PmmcError, match="job 0: schedule"):
        ctx.anneal_batch([ap, ap, ap], [bad, good, good], seeds=[1, 2, 3], n_chains=64)
    with pytest.raises(PmmcError, match="job 2: schedule"):
        ctx.anneal_batch([ap, ap, ap], [good, good, bad], seeds=[1, 2, 3], n_chains=64)
    nan = synth.arcs_problem(n=50, n_arcs=150, seed=73)
    nan.exp_dist = np.array(nan.exp_dist, np.float32)
    nan.exp_dist[3, 4] = np.nan
    with pytest.raises(PmmcError, match="job 1: arcs: NaN"):
        ctx.anneal_batch([ap, nan, nan], [good, good, good], seeds=[1, 2, 3], n_chains=64)
    # the context still works afterwards
    a = ctx.anneal_batch([ap, ap], [good, good], seeds=[1, 2], n_chains=64)
    b = ctx.anneal_batch([ap, ap], [good, good], seeds=[1, 2], n_chains=64)
    assert all(np.array_equal(x[0], y[0]) for x, y in zip(a, b))


def test_server_budget_admission(ctx):
    sched = schedule_preset(2, 1, settings(steps=1024), 0.3)
    sched.max_moves = 1024
    probs = [synth.arcs_problem(n=60, n_arcs=180, seed=80 + i) for i in range(4)]
    srv = ctx.server(hbm_budget=1 << 30, max_jobs=1, blocks_per_step=1)
    ids = [srv.submit(p, sched, seed=3, segment=i, n_chains=64) for i, p in enumerate(probs)]
    seen = []
    while not all(srv.done(j) for j in ids):
        _, r, q = srv.step()
        seen.append((r, q))
    assert max(r for r, _ in seen) <= 1          # one slot: jobs run one after another
    with pytest.raises(Exception):
        srv2 = ctx.server(hbm_budget=1024, max_jobs=4)
        srv2.submit(probs[0], sched, n_chains=64)   # larger than the budget
    for bad in (dict(streams=17), dict(streams=-1), dict(launch_ms=-1.0), dict(launch_ms=2000.0)):
        with pytest.raises(Exception, match="launch_ms"):
            ctx.server(hbm_budget=1 << 20, max_jobs=1, blocks_per_step=0, **bad)
    srv.close()


# ---------------- lane-pair arc kernel and the launch-shape tuner ----------------
def _pair_cases():
    s = settings(max_temp=5.0, dt_temp=0.9999, improvement=0.975, steps=2048, min_successes=100)
    probs, scheds = [], []
    for n, seed in ((30, 1), (150, 2), (400, 3)):
        p = synth.arcs_problem(n=n, n_arcs=3 * n, seed=90 + seed)
        probs.append(p)
        scheds.append(schedule_preset(2, 1, s, 0.3))
    # a hub bead with 40 arcs (rows longer than one fetched chunk), loops and
    # the sphere penalty, 2D moves
    hub = synth.arcs_problem(n=120, n_arcs=360, seed=97)
    ed = hub.exp_dist.copy()
    for j in range(1, 41):
        ed[0, j] = ed[j, 0] = 1.0 + 0.05 * j
    hub.exp_dist = ed
    hub.loop_pairs = np.array([[0, 60], [5, 100]], np.int32)
    hub.loop_params = np.array([[1.0, 1.5], [0.5, 3.0]], np.float32)
    sh = schedule_preset(2, 1, settings(max_temp=5.0, dt_temp=0.9999, improvement=0.975,
                                        steps=2048, min_successes=100, sphere_radius=6.0,
                                        use2d=1), 0.3)
    probs.append(hub)
    scheds.append(sh)
    return probs, scheds


@pytest.mark.parametrize("cpw", [32, 64])
def test_batch_arcs_lane_pair_equals_plain(ctx, cpw):
    probs, scheds = _pair_cases()
    chains = [100, 256, 64, 192]          # ragged: a partial last ensemble
    print(f"arcs batch chains_per_wave={cpw}", flush=True)
    alone = [ctx.anneal(p, sc, seed=21, segment=i, n_chains=chains[i])
             for i, (p, sc) in enumerate(zip(probs, scheds))]
    got = {}
    mixed = ctx.anneal_batch(probs, scheds, seeds=[21] * 4, n_chains=chains, blocks_per_step=3,
                             steps_per_graph=4, use_graph=True, chains_per_wave=cpw, opts_out=got)
    assert got["chains_per_wave"] == cpw and got["tuned"] == 0
    for (pa, sa), (pb, sb) in zip(alone, mixed):
        assert np.array_equal(pa, pb)
        for k in ("energy", "energy_resummed", "best_energy", "moves", "accepted", "best_chain"):
            assert sa[k] == sb[k], k


def test_batch_arcs_lane_pair_matches_oracle(ctx):
    # the lane-pair kernel's best chain against the oracle's chains
    probs, scheds = _pair_cases()
    p, sc = probs[3], scheds[3]
    sc.max_moves = 2048
    (pos, st), = ctx.anneal_batch([p], [sc], seeds=[7], segments=[3], n_chains=64,
                                  blocks_per_step=2, use_graph=False, chains_per_wave=32)
    P = O.Problem(p)
    res = [P.anneal(O.schedule_from(sc), 7, 3, c) for c in range(64)]
    best = int(np.argmin([r[1]["energy"] for r in res]))
    assert st["best_chain"] == best
    assert np.array_equal(pos, res[best][0])
    assert st["moves"] == sum(r[1]["moves"] for r in res)


def test_batch_tuner_chooses_and_keeps_results(ctx):
    probs, scheds = _pair_cases()
    chains = [64, 64, 128, 64]
    alone = [ctx.anneal(p, sc, seed=4, segment=i, n_chains=chains[i])
             for i, (p, sc) in enumerate(zip(probs, scheds))]
    got = {}
    mixed = ctx.anneal_batch(probs, scheds, seeds=[4] * 4, n_chains=chains, use_graph=True,
                             opts_out=got)
    print(f"tuner chose {got}", flush=True)
    assert got["tuned"] == 1 and got["chains_per_wave"] in (32, 64)
    assert got["blocks_per_step"] >= 1 and 1 <= got["steps_per_graph"] <= 64
    for (pa, sa), (pb, sb) in zip(alone, mixed):
        assert np.array_equal(pa, pb) and sa["energy"] == sb["energy"]


# ---------------- full-length chains at the workload sizes ----------------
# The trajectory tests above cap chains at a few thousand moves.  Here chains
# of the 64-bead heatmap workload (MonteCarloHeatmap's schedule, ~0.7-1 M
# moves each) and of the 400-anchor arc segment (MonteCarloArcs' schedule,
# 450 k moves each) run to their stop rule as single-chain jobs of one batch,
# and each is compared with the oracle's chain.  A chain whose oracle run
# made no decision within FP32 error of its threshold (near tie, tie_rel
# 1e-6) must end bit for bit where the oracle's does; one that did may part
# after its first near tie.  Long chains nearly all meet one (first after
# 4-8 k moves), so beyond that: at the heatmap level, where a near tie rarely
# flips, at least two thirds of the chains must still end bit for bit (12 of
# 12 measured); at the arc level, whose late moves sit at ~1e-6 relative
# margins (2 of 12 measured), the chains' final energies must agree with the
# oracle's within the band the reference comparison uses (|mean log-ratio|
# <= 1e-3).  The smooth level's chains (a 1,001-bead segment, 8.4 M moves to
# MonteCarloArcsSmooth's stop rule) meet no near tie and must all end bit for
# bit.
@pytest.mark.timeout(600)
@pytest.mark.parametrize("level,k", [(1, 12), (2, 12), (3, 8)])
def test_full_length_chains_match_oracle(ctx, level, k):
    # smooth: a 1,001-bead segment (21 anchors x 49), whose chains run 8.4 M
    # moves to MonteCarloArcsSmooth's stop rule without a near tie
    prob, step, _ = synth.config_case(level, 2, 21, 49) if level == 3 else synth.config_case(level, 2)
    sched = schedule_preset(level, 0, settings(**synth.CONFIG_SETTINGS[level]), step)
    chain0s = [64 * c for c in range(k)]
    print(f"full-length chains level {level}: {k} single-chain jobs", flush=True)
    res = ctx.anneal_batch([prob] * k, sched, seeds=[5] * k, segments=[9] * k,
                           chain0s=chain0s, n_chains=1)
    P = O.Problem(prob)
    osched = O.schedule_from(sched)
    import concurrent.futures as cf
    with cf.ThreadPoolExecutor(max_workers=8) as ex:   # the oracle's C calls release the GIL
        orc = list(ex.map(lambda c0: P.anneal(osched, 5, 9, c0, tie_rel=1e-6), chain0s))
    tie_free = exact = 0
    for c0, (pos, st), (opos, ost) in zip(chain0s, res, orc):
        same = (st["moves"] == ost["moves"] and st["accepted"] == ost["accepted"]
                and np.array_equal(pos, opos))
        exact += same
        if ost["near_ties"] == 0:
            tie_free += 1
            assert same, f"chain {c0}: no near tie in the oracle's run, yet the chains differ"
    lr = float(np.mean(np.log([st["energy"] / ost["energy"] for (_, st), (_, ost) in zip(res, orc)])))
    print(f"level {level}: {exact} of {k} chains bit-exact to their ends "
          f"({tie_free} without a near tie), "
          f"{np.mean([r[1]['moves'] for r in res]):.0f} moves per chain on average, "
          f"energy mean log-ratio {lr:.2e}", flush=True)
    if level == 1:
        assert exact >= (2 * k) // 3
    if level == 3:
        assert tie_free >= k // 2
    assert abs(lr) <= 1e-3


def test_headline_chains_match_oracle(ctx):
    """The headline workload as benchmarked: the 400-anchor segment under the
    GPU-kernel schedule of default.ini (chains capped at 50,176 moves).
    Twelve chains as single-chain jobs against the oracle's: every chain
    without a near tie bit for bit, and at least two thirds of all."""
    import concurrent.futures as cf
    import bench
    prob, s, step, _ = bench.workload()
    sched = schedule_preset(2, 1, s, step)
    k = 12
    chain0s = [64 * c for c in range(k)]
    res = ctx.anneal_batch([prob] * k, sched, seeds=[100] * k, segments=[0] * k,
                           chain0s=chain0s, n_chains=1)
    P = O.Problem(prob)
    osched = O.schedule_from(sched)
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        orc = list(ex.map(lambda c0: P.anneal(osched, 100, 0, c0, tie_rel=1e-6), chain0s))
    exact = tie_free = 0
    for c0, (pos, st), (opos, ost) in zip(chain0s, res, orc):
        same = (st["moves"] == ost["moves"] and st["accepted"] == ost["accepted"]
                and np.array_equal(pos, opos))
        exact += same
        if ost["near_ties"] == 0:
            tie_free += 1
            assert same, c0
    print(f"headline chains: {exact} of {k} bit-exact ({tie_free} without a near tie), "
          f"{np.mean([r[1]['moves'] for r in res]):.0f} moves each", flush=True)
    assert exact >= (2 * k) // 3
