"""GPU checks of context-level behaviour of the C ABI: the default chain count
does not depend on which problems a process ran before, the arena can be
reserved ahead of a call, and heatmap tables whose chromosome boundaries do
not partition the beads are evaluated with the general (asymmetric) form."""
import numpy as np
import pytest

from oracle import binding as O
from pmmc_amd import synth
from pmmc_amd.api import PmmcError, schedule_preset, settings

from test_gpu_parity import _check_delta

pytestmark = pytest.mark.gpu


def _asym(prob, seed):
    rng = np.random.default_rng(seed)
    up = np.triu(prob.dist > 0, 1)
    prob.dist = np.where(up, prob.dist * rng.uniform(0.8, 1.2, prob.dist.shape),
                         prob.dist).astype(np.float32)
    return prob


def test_default_chains_independent_of_call_order(ctx):
    # 33..64 beads: symmetric tables run the lane-pair kernel, asymmetric ones
    # the one-lane-per-chain kernel; their occupancies differ, and the default
    # chain count of one must not depend on whether the other ran first
    sym, _ = synth.heatmap_problem(n=48, seed=21)
    asy = _asym(synth.heatmap_problem(n=50, seed=22)[0], 22)
    sched = schedule_preset(1, 1, settings(steps=512), 1.0)
    sched.max_moves = 512
    print("default chains: symmetric 48 / asymmetric 50 / symmetric 48", flush=True)
    _, s1 = ctx.anneal(sym, sched, seed=3, n_chains=0)
    _, s2 = ctx.anneal(asy, sched, seed=3, n_chains=0)
    _, s3 = ctx.anneal(sym, sched, seed=3, n_chains=0)
    assert s1["n_chains"] == s3["n_chains"] == ctx.default_chains(1, 48, True)
    assert s2["n_chains"] == ctx.default_chains(1, 50, False)
    assert s1["n_chains"] % 64 == 0 and s2["n_chains"] % 64 == 0
    # at least 16 wavefronts per compute unit (one full round or more)
    assert s1["n_chains"] >= 16 * 64 * ctx.cu_count


def test_default_chains_levels(ctx):
    for level, n in ((2, 400), (3, 19951), (1, 24), (1, 200)):
        c = ctx.default_chains(level, n)
        assert c > 0 and c % 64 == 0, (level, n, c)
    assert ctx.default_chains(4, 10) == 0 and ctx.default_chains(2, 0) == 0


def test_delta_heatmap_boundaries_not_partitioning(ctx):
    # boundaries {0, 60} with n = 150: beads 0..59 see the block [0, 60),
    # beads 60.. the whole range, so term j of local(p) and term p of local(j)
    # differ although the targets are symmetric
    prob, _ = synth.heatmap_problem(n=150, seed=2, diagonal=2)
    prob.boundaries = np.array([0, 60], np.int32)
    print("delta heatmap n=150 boundaries [0, 60]", flush=True)
    _check_delta(ctx, prob)
    _check_delta(ctx, prob, r=4.0, seed=1)
    # and a lane-pair-sized one (33..64 beads), boundaries {0, 20}
    prob, _ = synth.heatmap_problem(n=48, seed=3)
    prob.boundaries = np.array([0, 20], np.int32)
    _check_delta(ctx, prob)


def test_trajectory_boundaries_not_partitioning(ctx):
    prob, _ = synth.heatmap_problem(n=48, seed=3)
    prob.boundaries = np.array([0, 20], np.int32)
    s = settings(max_temp=5.0, dt_temp=0.999, improvement=0.99, steps=512, min_successes=10,
                 milestone_fails=3)
    sched = schedule_preset(1, 1, s, 1.0)
    sched.max_moves = 2048
    print("trajectory heatmap n=48 boundaries [0, 20]", flush=True)
    pos, st = ctx.anneal(prob, sched, seed=7, segment=3, chain0=0, n_chains=64)
    P = O.Problem(prob)
    osched = O.schedule_from(sched)
    res = [P.anneal(osched, 7, 3, c) for c in range(64)]
    best = int(np.argmin([r[1]["energy"] for r in res]))
    assert st["best_chain"] == best
    assert np.array_equal(pos, res[best][0])
    assert st["energy_resummed"] == pytest.approx(P.total(pos), rel=1e-12)


def test_ctx_reserve(ctx):
    print("ctx reserve 1 GiB, then an anneal", flush=True)
    ctx.reserve(1 << 30)
    prob = synth.arcs_problem(n=60, n_arcs=180, seed=4)
    sched = schedule_preset(2, 1, settings(steps=1024), 0.3)
    sched.max_moves = 1024
    a, sa = ctx.anneal(prob, sched, seed=2, n_chains=128)
    # a reservation larger than the device fails cleanly and leaves the
    # context usable
    with pytest.raises(PmmcError):
        ctx.reserve(1 << 50)
    b, sb = ctx.anneal(prob, sched, seed=2, n_chains=128)
    assert np.array_equal(a, b) and sa["energy"] == sb["energy"]
