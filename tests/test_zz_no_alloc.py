"""The allocation-free annealing entry point (pmmc_anneal_noalloc) against the
plain entry points: the same kernels on buffers carved from the library's
static per-device scratch, so results are bit-identical.  Two calls in a
row, the second with a smaller input (the scratch keeps the first call's
contents), per level; the default chain count that fits the scratch; and a
request larger than the scratch, rejected with a clear error."""
import numpy as np
import pytest

from pmmc_amd import synth
from pmmc_amd.api import PmmcError, schedule_preset, settings

pytestmark = pytest.mark.gpu


def _same(a, b):
    (pa, sa), (pb, sb) = a, b
    assert np.array_equal(pa, pb)
    for k in ("energy", "energy_resummed", "best_energy", "moves", "accepted", "best_chain",
              "best_chain_moves", "n_chains", "milestones", "stop_reason"):
        assert sa[k] == sb[k], k


def test_noalloc_arcs_two_calls(ctx):
    s = settings(max_temp=5.0, dt_temp=0.9999, improvement=0.975, steps=4096, min_successes=100)
    sched = schedule_preset(2, 1, s, 0.3)
    big = synth.arcs_problem(n=400, n_arcs=1200, seed=61)
    small = synth.arcs_problem(n=90, n_arcs=270, seed=62)
    print("noalloc arcs n=400 4096 chains, then n=90 1024 chains", flush=True)
    a1 = ctx.anneal_noalloc(big, sched, seed=5, segment=2, n_chains=4096)
    a2 = ctx.anneal_noalloc(small, sched, seed=5, segment=3, n_chains=1024)
    _same(a1, ctx.anneal(big, sched, seed=5, segment=2, n_chains=4096))
    _same(a2, ctx.anneal(small, sched, seed=5, segment=3, n_chains=1024))


def test_noalloc_heatmap_and_smooth_two_calls(ctx):
    hs = schedule_preset(1, 1, settings(steps=512), 1.0)
    hs.max_moves = 4096
    h1, _ = synth.heatmap_problem(n=64, seed=7)
    h2, _ = synth.heatmap_problem(n=20, seed=8)
    print("noalloc heatmap n=64 then n=20", flush=True)
    r1 = ctx.anneal_noalloc(h1, hs, seed=9, n_chains=2048)
    r2 = ctx.anneal_noalloc(h2, hs, seed=9, n_chains=640)
    _same(r1, ctx.anneal(h1, hs, seed=9, n_chains=2048))
    _same(r2, ctx.anneal(h2, hs, seed=9, n_chains=640))
    ss = schedule_preset(3, 1, settings(steps=4096, dt_temp=0.9995), 0.5)
    s1 = synth.smooth_problem(n_anchors=60, add=9, seed=3)
    s2 = synth.smooth_problem(n_anchors=10, add=9, seed=4)
    print("noalloc smooth 60 anchors then 10 anchors", flush=True)
    q1 = ctx.anneal_noalloc(s1, ss, seed=2, n_chains=1024)
    q2 = ctx.anneal_noalloc(s2, ss, seed=2, n_chains=256)
    _same(q1, ctx.anneal(s1, ss, seed=2, n_chains=1024))
    _same(q2, ctx.anneal(s2, ss, seed=2, n_chains=256))
    # fixed beads keep the caller's positions (smooth writes back free beads only)
    assert np.array_equal(q1[0][s1.fixed == 1], s1.pos[s1.fixed == 1])


def test_noalloc_default_chains_fit_and_limit(ctx):
    s = settings(max_temp=5.0, dt_temp=0.9999, improvement=0.975, steps=1024, min_successes=100)
    sched = schedule_preset(2, 1, s, 0.3)
    sched.max_moves = 1024
    prob = synth.arcs_problem(n=400, n_arcs=1200, seed=63)
    print("noalloc arcs n=400 default chains", flush=True)
    p, st = ctx.anneal_noalloc(prob, sched, seed=4, n_chains=0)
    per_ens = 400 * 3 * 64 * 4
    assert st["n_chains"] % 64 == 0 and 0 < st["n_chains"] * per_ens // 64 <= 64 << 20
    assert st["n_chains"] >= 64 * ((48 << 20) // per_ens)     # close to what fits
    _same((p, st), ctx.anneal(prob, sched, seed=4, n_chains=st["n_chains"]))
    print("noalloc arcs n=400 262144 chains: rejected", flush=True)
    with pytest.raises(PmmcError) as e:
        ctx.anneal_noalloc(prob, sched, seed=4, n_chains=262144)
    assert e.value.status == 3 and "scratch" in str(e.value)
    # the entry point stays usable after the rejection
    q = ctx.anneal_noalloc(prob, sched, seed=4, n_chains=st["n_chains"])
    _same(q, (p, st))
