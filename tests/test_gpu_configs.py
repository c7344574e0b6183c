"""The batched and served forms at their workloads' sizes, with capped move
counts to keep the GPU time short.

Batched (512 arc problems in one pool: 32 segments of 20-1,974 anchors x 16
ensemble seeds): every problem's result is the plain call's, bit for bit, on
8 problems including the largest; two problems against the oracle's chains.

Served (a Poisson-like mix at the workload's sizes: heatmaps of 64-512
beads, arc problems of 20-2,000 anchors, smooth chains of 1,001-19,951 beads,
1-3 solves per request): jobs are submitted while others run, and every
solve equals its plain call, bit for bit.
"""
import numpy as np
import pytest

from oracle import binding as O
from pmmc_amd import schedule_preset, settings, synth

pytestmark = pytest.mark.gpu

TIE_REL = 1e-6
TIE = 1e-5


def _assert_oracle_best(prob, sched, seed, segment, pos, st):
    """The batched job's returned chain against the oracle's chains (same
    Philox streams): bit-exact unless a decision of that chain was a near tie;
    its energy is the lowest up to near ties."""
    P = O.Problem(prob)
    n_chains = st["n_chains"]
    out, stats, _ = P.anneal_many(O.schedule_from(sched), seed, segment, 0, n_chains, 8,
                                  tie_rel=TIE_REL)
    E = np.array([s["energy"] for s in stats])
    gb = st["best_chain"]
    if stats[gb]["near_ties"] == 0:
        assert np.array_equal(pos, out[gb])
        assert st["energy"] == pytest.approx(E[gb], rel=1e-6)
        assert st["best_chain_moves"] == stats[gb]["moves"]
    scale = abs(P.total()) + abs(E.min())
    assert E[gb] <= E.min() + TIE * scale or any(s["near_ties"] for s in stats)
    return int(np.argmin(E)) == gb


def test_config3_batch_512_problems(ctx):
    probs, meta = synth.batch_arcs_problems(n_segments=32, n_seeds=16, seed=0)
    assert len(probs) == 512
    sizes = [m[2] for m in meta]
    big = int(np.argmax(sizes))
    assert sizes[big] == 1974
    s = settings(max_temp=5.0, dt_temp=0.9999, improvement=0.975, steps=1024, min_successes=100)
    sched = schedule_preset(2, 1, s, 0.03)
    # two small problems run 64 chains each for the oracle comparison
    small = [i for i in np.argsort(sizes)[:40:20]]
    chains = [64 if i in small else 128 for i in range(512)]
    seeds = [1 + i for i in range(512)]
    print(f"config 3: 512 problems, anchors {min(sizes)}-{max(sizes)}", flush=True)
    res = ctx.anneal_batch(probs, sched, seeds=seeds, n_chains=chains, use_graph=True)
    assert all(r[1]["moves"] > 0 for r in res)
    check = sorted(set([big] + small + list(range(0, 512, 97))))[:8]
    assert big in check
    for i in check:
        print(f"plain problem {i} ({sizes[i]} anchors)", flush=True)
        pos, st = ctx.anneal(probs[i], sched, seed=seeds[i], segment=i, n_chains=chains[i])
        assert np.array_equal(pos, res[i][0]), i
        for k in ("energy", "energy_resummed", "moves", "accepted", "best_chain", "milestones"):
            assert st[k] == res[i][1][k], (i, k)
    for i in small:
        print(f"oracle problem {i} ({sizes[i]} anchors)", flush=True)
        _assert_oracle_best(probs[i], sched, seeds[i], i, *res[i])


def _served_jobs(rng):
    hs = schedule_preset(1, 1, settings(steps=50000), 1.0)
    hs.max_moves = 1024
    asched = schedule_preset(2, 1, settings(steps=1024, improvement=0.975, min_successes=100), 0.03)
    ss = schedule_preset(3, 1, settings(steps=1024, dt_temp=0.9995, improvement=0.99,
                                        min_successes=50), 0.3)
    jobs = []
    for n in (512, 64, 150, 333):
        jobs.append((synth.heatmap_problem(n=n, seed=n)[0], hs))
    for n in (2000, 20, 180, 900):
        jobs.append((synth.arcs_problem(n=n, n_arcs=3 * n, seed=n), asched))
    for a in (400, 21, 100):
        jobs.append((synth.smooth_problem(n_anchors=a, add=49, seed=a), ss))
    order = rng.permutation(len(jobs))
    return [jobs[k] + (int(rng.integers(1, 4)),) for k in order]


def test_config4_served_mix_workload_sizes(ctx):
    rng = np.random.default_rng(4)
    reqs = _served_jobs(rng)
    sizes = [len(np.asarray(p.pos).reshape(-1, 3)) for p, _, _ in reqs]
    print(f"config 4: {len(reqs)} requests, beads {min(sizes)}-{max(sizes)}", flush=True)
    assert max(sizes) == 19951
    srv = ctx.server(hbm_budget=16 << 30, max_jobs=64, blocks_per_step=0)
    ids = {}
    nxt, steps, admitted_while_busy, max_res = 0, 0, 0, 0
    resident = 0
    while nxt < len(reqs) or not all(srv.done(j) for v in ids.values() for j in v):
        # arrivals: 1-2 requests per step while any remain
        for _ in range(int(rng.integers(1, 3)) if steps else 3):
            if nxt < len(reqs):
                p, sc, ens = reqs[nxt]
                if resident > 0:
                    admitted_while_busy += 1
                ids[nxt] = [srv.submit(p, sc, seed=1 + e, segment=nxt, n_chains=64) for e in range(ens)]
                nxt += 1
        _, resident, _ = srv.step()
        max_res = max(max_res, resident)
        steps += 1
        assert steps < 100000
    print(f"served in {steps} steps, up to {max_res} resident jobs, "
          f"{admitted_while_busy} requests submitted while others ran", flush=True)
    assert admitted_while_busy >= 3 and max_res >= 2
    for k, (p, sc, ens) in enumerate(reqs):
        for e in range(ens):
            pos, st = ctx.anneal(p, sc, seed=1 + e, segment=k, n_chains=64)
            spos, sst = srv.result(ids[k][e])
            assert np.array_equal(pos, spos), (k, e, sizes[k])
            for key in ("energy", "energy_resummed", "moves", "best_chain", "milestones"):
                assert st[key] == sst[key], (k, e, key)
    srv.close()
