"""Benchmark of the pMMC annealing hot path on one MI355X.

Headline (`value`): plain arc-level simulated annealing of one 2 Mb segment
(400 anchors, 1,200 PET-cluster arcs).  One step = one plain solve
(pmmc_arcs_anneal: every chain runs to its stop rule, the best chain is
returned; inputs are uploaded inside the step, as the drop-in does).  `value`
is counted MC moves per second over the timed steps (whole job).  The
roofline figure uses the annealing kernel's own duration, measured with HIP
events on the library's stream.

Beside it, in the same line:
  batched -- 512 arc problems (32 segments x 16 ensemble seeds, 20-2,000
             anchors) in one device-resident pool with a captured step loop,
             against the same problems solved one plain call at a time
             (timed on a subset, per-problem results checked identical);
  served  -- Poisson arrivals of mixed heatmap / arc / smooth jobs through the
             continuous-batching server, p50/p99 request latency and solves/s,
             against a sample of the same requests solved one plain call at a
             time (results checked identical);
  levels  -- the plain heatmap-level workload (64 beads, MonteCarloHeatmap's
             schedule) and the plain smooth-level workload (19,951 beads, the
             smooth GPU kernel's schedule): moves/s, final energy, and for the heatmap the
             distance-map Pearson / SCC / RMSD to the ground-truth polymer,
             each beside the reference's CPU annealer on the host's cores;
  matched_energy -- the headline segment annealed to completion under the
             reference CPU annealer's own schedule: complete anneals per
             second on the GPU against the CPU baseline's, with each side's
             median final energy of one chain (the headline's GPU-kernel
             schedule caps a chain at 50,176 moves, so its solves end at a
             higher energy than the CPU's complete anneals);
  cpu_baseline -- the reference's own CPU arc annealer (LooperSolver::
             MonteCarloArcs, compiled in place into oracle/_ref) timed on the
             host's cores; the oracle's restatement when that build is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MC moves/s + batched solves/s vs plain GPU; p50/p99 latency at matched energy"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_anneal_arcs_n400_hbm_traffic.txt")
PROBE_FILE = os.path.join(ROOT, "profiles", "probe_move_patterns.txt")


def genomic_step(anchors, noise=0.01, power=0.75, scale=0.5, base=1.0):
    """noise_size of LooperSolver::reconstructClusterArcsDistances
    (LooperSolver.hpp:4339-4354): mean genomicLengthToDistance of consecutive
    anchors times noise_arcs ([distance] and [main] of default.ini)."""
    import numpy as np
    g = anchors.mean(axis=1)
    d = base + scale * np.power(np.diff(g) / 1000.0, power)
    return float(d.sum() / len(anchors) * noise)


def arcs_settings():
    from pmmc_amd import settings
    # [simulation_arcs] and [cuda] of auxiliary/configs/default.ini
    return settings(max_temp=5.0, dt_temp=0.9999, improvement=0.975, tjs=50.0, tjc=20.0,
                    min_successes=100, steps=50000, milestone_fails=3)


def workload():
    from pmmc_amd import synth
    anchors, arcs, _ = synth.arcs_inputs(n=400, n_arcs=1200, seed=0)
    prob = synth.arcs_problem(n=400, n_arcs=1200, seed=0)
    return prob, arcs_settings(), genomic_step(anchors), arcs


def pattern_ceiling(key="arc_pattern_ceiling_moves_per_s"):
    """Moves/s the move kernel's own access pattern sustains with no
    arithmetic, from the committed bench.py --probe summary, or None."""
    try:
        for line in open(PROBE_FILE):
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_traffic_per_move():
    """HBM bytes per move of the arc kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by its copy-kernel
    calibration factor), or None when the summary is absent."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("hbm_bytes_per_move:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_calibration() -> int:
    """Three device-to-device copies of 1 GiB (torch's copy kernel): reading
    FETCH_SIZE / WRITE_SIZE for them gives the factor between each counter
    and bytes actually moved."""
    import torch
    n = 1 << 28
    src = torch.ones(n, device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copied_bytes_per_launch": n * 4, "launches": 3}))
    return 0


def host_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def cpu_reference(level, prob, step, settings_dict, seconds, sample_moves=0):
    """The reference's CPU annealer of `level` (compiled in place into
    oracle/_ref/levels_ref), one process per host thread, each annealing the
    workload repeatedly for about `seconds`.  sample_moves > 0 bounds each run
    to exactly that many moves (milestone period sample_moves with a stop test
    that always fires: improvement 0, min successes 2^31-1); the reference's
    per-move work does not depend on acceptance.  None when that build is
    absent."""
    from oracle import refdriver as R
    if not R.available():
        return None
    threads = host_threads()
    sd = dict(settings_dict)
    if sample_moves > 0:
        sd.update(steps=sample_moves, improvement=0.0, min_successes=2**31 - 1)
    moves, secs, runs, scores = R.time_anneal(level, prob, step, sd, seconds, threads)
    out = {"value": moves / secs, "unit": "moves/s", "cores": threads, "kind": "reference",
           "per_core": moves / secs / threads, "runs": runs, "seconds": secs,
           "sample": (f"{runs} runs of {sample_moves} moves" if sample_moves > 0 else
                      f"{runs} complete anneals") + f" on {threads} host threads"}
    if sample_moves == 0:
        # complete anneals: solves per second and the energy they reach
        out["runs_per_s"] = runs / secs
        out["final_energy_median"] = float(sorted(scores)[len(scores) // 2])
    return out


def cpu_baseline(prob, s, step, budget_s=15.0):
    """The reference's CPU arc annealer (LooperSolver::MonteCarloArcs) on the
    host's cores; the oracle's restatement of it when the reference build is
    absent."""
    from pmmc_amd import synth
    ref = cpu_reference(2, prob, step, synth.CONFIG_SETTINGS[2], budget_s)
    if ref is not None:
        ref["sample"] = (f"{ref['runs']} complete anneals of the 400-anchor segment by the "
                         f"reference's LooperSolver::MonteCarloArcs (default.ini "
                         f"[simulation_arcs], glibc rand), {ref['seconds']:.1f} s on each of "
                         f"{ref['cores']} host threads")
        return ref
    from oracle import binding as O
    from pmmc_amd import schedule_preset
    sched = O.schedule_from(schedule_preset(2, 0, s, step))
    P = O.Problem(prob)
    threads = host_threads()
    moves_per_chain = 50000
    sched.max_moves = moves_per_chain
    sched.block = 50000
    sched.milestone_period = 50000
    _, _, per_round = P.anneal_many(sched, 1, 0, 0, threads, threads)
    rounds = max(1, int(budget_s / max(per_round, 1e-3)))
    chains = threads * rounds
    _, stats, secs = P.anneal_many(sched, 1, 0, 0, chains, threads)
    moves = sum(x["moves"] for x in stats)
    return {"value": moves / secs, "unit": "moves/s", "cores": threads, "kind": "port",
            "per_core": moves / secs / threads,
            "sample": f"{chains} chains x {moves_per_chain} moves of the oracle's restatement "
                      f"of LooperSolver::MonteCarloArcs (dense O(n) local scans), "
                      f"{secs:.1f} s on {threads} host threads"}


def heatmap_leg(ctx, with_cpu=True, cpu_seconds=8.0):
    """Plain heatmap workload, 64 beads (init U[0,10)^3, target from a
    random-walk polymer), on the schedule of the CPU path the workload is
    specified on (LooperSolver::MonteCarloHeatmap: 50k-move milestones,
    cooling per move): one solve after a warm-up."""
    from pmmc_amd import metrics, schedule_preset, settings, synth
    prob, step, truth = synth.config_case(1, 0)
    sched = schedule_preset(1, 0, settings(**synth.CONFIG_SETTINGS[1]), step)
    # 4 chains per slot: one round of the lane-pair register kernel's workers
    # (6 per CU, 98,304 slots) takes 393,216 chains from the queue
    chains = 4 * 98304
    warm = schedule_preset(1, 0, settings(**synth.CONFIG_SETTINGS[1]), step)
    warm.max_moves = 1024
    ctx.anneal(prob, warm, seed=1, n_chains=chains)
    t0 = time.perf_counter()
    pos, st = ctx.anneal(prob, sched, seed=2, segment=1, n_chains=chains)
    wall = time.perf_counter() - t0
    h = {"workload": "plain heatmap-level SA, 64 beads, default.ini [simulation_heatmap], "
                     "schedule of LooperSolver::MonteCarloHeatmap", "chains": st["n_chains"],
         "moves_per_s": st["moves"] / wall, "accepted_moves_per_s": st["accepted"] / wall,
         "kernel_moves_per_s": st["moves"] / (st["kernel_ms"] * 1e-3), "solve_s": wall,
         "final_energy": st["energy_resummed"],
         "partner_evaluations_per_s": 2 * 63 * st["moves"] / (st["kernel_ms"] * 1e-3),
         "bound": "issue latency (a chain's 33-64 bead positions held by a lane pair, 96 VGPRs "
                  "each: 3 waves per SIMD; a slot whose chain stops takes the next chain of the "
                  "queue; 12 B read + 12 B x acceptance written per chain-move)"}
    h.update(metrics.compare(pos, truth))
    if with_cpu:
        h["cpu_reference"] = cpu_reference(1, prob, step, synth.CONFIG_SETTINGS[1], cpu_seconds,
                                           sample_moves=20000)
    return h


def smooth_leg(ctx, with_cpu=True, cpu_seconds=8.0):
    """Plain smooth workload, 19,951 beads (400 fixed anchors, 49 beads
    between neighbours), on the GPU kernel's schedule: one solve after a
    warm-up solve."""
    from pmmc_amd import schedule_preset, settings, synth
    prob, step, _ = synth.config_case(3, 0)
    sm = dict(synth.CONFIG_SETTINGS[3])
    sched = schedule_preset(3, 1, settings(**sm), step)
    ctx.anneal(prob, sched, seed=1)
    t0 = time.perf_counter()
    pos, st = ctx.anneal(prob, sched, seed=2, segment=1)
    wall = time.perf_counter() - t0
    acc = st["accepted"] / max(1, st["moves"])
    # per chain-move: the 5-bead window (12 B each), the two rest lengths and
    # the free-list entry of the pick (12 B, one scalar read shared by the 64
    # chains of an ensemble) and the 12 B write of an accepted move
    bpm = 12.0 * 5 + 12.0 / 64 + 12.0 * acc
    kms = st["kernel_ms"] * 1e-3
    out = {
        "workload": "plain smooth-level SA, 19,951 beads, 200k-move cap, cooling 0.9995, "
                    "GPU-kernel schedule", "chains": st["n_chains"],
        "moves_per_s": st["moves"] / wall, "accepted_moves_per_s": st["accepted"] / wall,
        "kernel_moves_per_s": st["moves"] / kms, "solve_s": wall,
        "final_energy": st["energy_resummed"],
        "bytes_per_move": bpm, "achieved_GBps": bpm * st["moves"] / kms / 1e9,
        "frac_of_pattern_ceiling": (None if pattern_ceiling("smooth_pattern_ceiling_moves_per_s") is None
                                    else st["moves"] / kms / pattern_ceiling("smooth_pattern_ceiling_moves_per_s")),
        "bound": "VALU and HBM together (~425 VALU instructions per 64-chain move; "
                 "3.8 KB window read per 64-chain move)"}
    if with_cpu:
        out["cpu_reference"] = cpu_reference(3, prob, step, sm, cpu_seconds,
                                             sample_moves=2000000)
    return out


def heatmap_sweep_leg(ctx, with_cpu=True, cpu_seconds=8.0):
    """The reference GPU heatmap kernel's own algorithm (pmmc_heatmap_sweep)
    on the 64-bead heatmap, launched as default.ini sizes that kernel:
    blocks_multiplier 16 x 64 beads x 512 threads = 8,192 proposers per bead,
    512 proposals each per iteration, step 0.75 x step_size, 3 milestone
    fails ([cuda], [simulation_heatmap]).  One solve after a one-iteration
    warm-up."""
    from pmmc_amd import _abi as A
    from pmmc_amd import metrics, synth
    del with_cpu, cpu_seconds  # no CPU counterpart: the reference runs it on CUDA only
    prob, step, truth = synth.config_case(1, 0)
    cs = synth.CONFIG_SETTINGS[1]
    # blocks_multiplier x n blocks of num_threads: 16 x 512 proposers per bead
    proposers, inner = 16 * 512, 512
    ss = A.SweepSchedule(cs["max_temp"], cs["dt_temp"], cs["improvement"], 1e-4, cs["tjs"],
                         cs["tjc"], 0.75 * step, 0.95, 0.0, inner, proposers, 3, 1000, 0)
    warm = A.SweepSchedule(cs["max_temp"], cs["dt_temp"], cs["improvement"], 1e-4, cs["tjs"],
                           cs["tjc"], 0.75 * step, 0.95, 0.0, inner, proposers, 3, 1, 0)
    ctx.heatmap_sweep(prob, warm, seed=1, segment=0)
    t0 = time.perf_counter()
    pos, st = ctx.heatmap_sweep(prob, ss, seed=2, segment=1)
    wall = time.perf_counter() - t0
    proposals = st["moves"]  # iterations x free beads x proposers x inner
    kern = st["kernel_ms"] * 1e-3
    out = {"workload": "reference GPU heatmap algorithm (parallel proposers, race-free sweep), "
                       "64 beads, 8,192 proposers x 512 proposals per bead per iteration",
           "iterations": st["milestones"], "proposals": proposals, "solve_s": wall,
           "proposals_per_s": proposals / wall, "kernel_proposals_per_s": proposals / kern,
           "partner_evaluations_per_s": proposals * (len(prob.pos) - 1) / kern,
           "final_energy": st["energy"], "stop_reason": st["stop_reason"],
           "bound": "VALU (snapshot and weight column in LDS)"}
    out.update(metrics.compare(pos, truth))
    return out


def singleton_leg(ctx, with_cpu=True, cpu_seconds=8.0):
    """createSingletonHeatmap on the GPU (pmmc_singleton_heatmap): 20 M
    singleton records (70% intra-chromosome) binned onto the anchor level of
    3 chromosomes x 200 anchors (600 bins: sums in the device matrix) and onto
    12 segment-like bins (sums in per-workgroup LDS histograms).  Kernel time
    from HIP events around the binning launch; end to end includes the
    records' upload and the host's O(bins^2) float steps."""
    import numpy as np
    del cpu_seconds
    rng = np.random.default_rng(11)
    out = {}
    for name, per_chr in (("anchor_600_bins", 200), ("segment_12_bins", 4)):
        chrs, clusters = [], []
        for c in range(3):
            gaps = rng.integers(2000, 30000, size=per_chr)
            lens = rng.integers(500, 8000, size=per_chr)
            starts = 50_000 + np.cumsum(gaps + lens) - lens
            chrs.append(f"chr{c + 1}")
            clusters.append([(int(a), int(a + b)) for a, b in zip(starts, lens)])
        n = 20_000_000
        hi = np.array([cl[-1][1] + 20_000 for cl in clusters], np.int64)
        c1 = rng.integers(0, 3, size=n)
        c2 = np.where(rng.random(n) < 0.7, c1, rng.integers(0, 3, size=n))
        s1 = (rng.random(n) * hi[c1]).astype(np.int64)
        s2 = (rng.random(n) * hi[c2]).astype(np.int64)
        rec = {k: v.astype(np.int32) for k, v in dict(
            chr1=c1, start1=s1, end1=s1 + 150, chr2=c2, start2=s2, end2=s2 + 150,
            score=rng.integers(1, 5, size=n)).items()}
        ctx.singleton_heatmap(chrs, clusters, rec)
        t0 = time.perf_counter()
        _, cnt = ctx.singleton_heatmap(chrs, clusters, rec)
        wall = time.perf_counter() - t0
        ks = cnt["kernel_ms"] * 1e-3
        leg = {"workload": f"singleton records binned onto {3 * per_chr} bins of 3 chromosomes",
               "records": n, "kernel_records_per_s": n / ks, "records_per_s": n / wall,
               "kernel_ms": cnt["kernel_ms"], "added": cnt["added"],
               "bytes_per_record": 28, "achieved_GBps": 28 * n / ks / 1e9,
               "frac_of_hbm_peak": 28 * n / ks / 1e9 / HBM_PEAK_GBS,
               "bound": ("atomics on the device matrix" if per_chr > 20 else
                         "LDS histogram atomics")}
        if with_cpu and name == "anchor_600_bins":
            from oracle import refdriver
            if refdriver.available():
                m = 2_000_000
                sub = {k: v[:m] for k, v in rec.items()}
                secs, cnt_ref = refdriver.singleton_time(chrs, clusters, sub)
                leg["cpu_reference"] = {
                    "value": cnt_ref / secs, "unit": "records/s", "cores": 1, "kind": "reference",
                    "sample": f"LooperSolver::createSingletonHeatmap on {cnt_ref} of the records "
                              "(the same bins), one thread, reading them from a singleton text "
                              "file as the reference does"}
        out[name] = leg
    return out


def stress_case(n, seed):
    """One chromosome of n anchors 4-30 kb apart, arcs spanning up to 6
    anchors with PET counts 2-6, n/4 long arcs, two segments, and a random
    walk of start positions."""
    import numpy as np
    rng = np.random.default_rng(seed)
    gaps = rng.integers(4000, 30000, size=n)
    lens = rng.integers(500, 3000, size=n)
    starts = 10_000 + np.cumsum(gaps + lens) - lens
    anchors = [(int(a), int(a + b)) for a, b in zip(starts, lens)]
    pairs = sorted({(int(a), int(min(n - 1, a + rng.integers(1, 7))))
                    for a in rng.integers(0, n - 1, size=n)})
    arcs = [(a, b, int(rng.integers(2, 7))) for a, b in pairs]
    longs = []
    for _ in range(n // 4):
        a, b = sorted(int(x) for x in rng.integers(0, n, size=2))
        longs.append((int(rng.integers(anchors[a][0], anchors[a][1] + 1)),
                      int(rng.integers(anchors[b][0], anchors[b][1] + 1)), int(rng.integers(2, 7))))
    lo, hi = anchors[0][0] - 1000, anchors[-1][1] + 1000
    cut = int(rng.integers(lo + 1, hi))
    case = dict(anchors=anchors, arcs=arcs, long_arcs=longs,
                splits=[("chr1", lo, cut), ("chr1", cut, hi)])
    pos = np.cumsum(rng.normal(0, 2.0, size=(n, 3)), 0).astype(np.float32)
    return case, pos


def arcs_level_leg(ctx, with_cpu=True, cpu_seconds=8.0):
    """A whole arc level (pmmc_arcs_level) over a synthetic genome: 4
    chromosomes x 600 anchors in 82 interaction blocks, default.ini's driver
    settings (one restart per block and stage, loop_density 5, noise 0.01 /
    5.0, [simulation_arcs] / [simulation_arcs_smooth] on the GPU-kernel
    presets), 256 chains per restart.  The driver's batched form (every
    block of a stage in one launch) against the same jobs issued one plain
    call each, in order, as the reference's reconstructClustersArcsDistances
    dispatches them (LooperSolver.hpp:4027-4138, :4408-4463); both give the
    same tree bit for bit.  Wall time includes the host's job preparation and
    densification."""
    import numpy as np
    from pmmc_amd import settings, synth
    from pmmc_amd import tree as T
    g = synth.genome(seed=5, n_chr=4, n_anchors=600)
    p = T.level_params(loop_density=5, noise_anchor=0.01, noise_subanchor=5.0,
                       restarts_anchor=1, restarts_subanchor=1)
    s_arcs = arcs_settings()
    s_smooth = settings(max_temp=5.0, dt_temp=0.9999, improvement=0.99, tjs=50.0, tjc=20.0,
                        min_successes=50, steps=50000, milestone_fails=3)

    def run(one_call):
        t = T.Tree(g["chrs"], g["splits"])
        rng = np.random.default_rng(7)
        for c in range(len(g["chrs"])):
            ibs = t.level(c, 2)
            t.set_positions(ibs, np.cumsum(rng.normal(0, 3.0, size=(len(ibs), 3)), 0).astype(np.float32))
        po = T.arcs_level_opts(arcs=s_arcs, smooth=s_smooth, chains_per_restart=256,
                               one_call_per_job=one_call)
        t0 = time.perf_counter()
        rep = T.arcs_level(ctx, t, -1, p, po, seed=1, segment=0)
        return t, rep, time.perf_counter() - t0

    run(0)                                  # grows the device arena once
    tb, rb, wb = run(0)
    ts, rs, ws = run(1)
    n = len(tb)
    pb, ps = tb.positions(list(range(n))), ts.positions(list(range(n)))
    return {"workload": "arc level of a synthetic genome: 4 chromosomes x 600 anchors, "
                        "82 interaction blocks, loop_density 5, one restart per block and stage, "
                        "256 chains per restart, GPU-kernel presets of default.ini",
            "beads_after": n, "jobs": int(rb["n_jobs"]), "moves": int(rb["moves"]),
            "batched_s": wb, "batched_ms_anchor": rb["ms_anchor"], "batched_ms_smooth": rb["ms_smooth"],
            "one_call_per_job_s": ws, "one_call_ms_anchor": rs["ms_anchor"],
            "one_call_ms_smooth": rs["ms_smooth"], "speedup": ws / wb,
            "batched_moves_per_s": rb["moves"] / wb,
            "identical": bool(np.array_equal(pb.view(np.uint32), ps.view(np.uint32))
                              and rb["score_anchor"] == rs["score_anchor"]
                              and rb["score_smooth"] == rs["score_smooth"])}


def stress_leg(ctx, with_cpu=True, cpu_seconds=8.0):
    """LooperSolver::stressRefineAnchors (the global anchor refinement's
    gradient descent) on one chromosome of 20,000 anchors: pmmc_tree_stress_
    refine (set-up on the host, every iteration in one cooperative launch) vs
    the reference's own function on one host thread, same inputs; the
    positions are compared bit for bit."""
    import numpy as np
    from pmmc_amd import tree as T
    del cpu_seconds
    n, iters, lr, ks, kq, ka = 20000, 300, 0.01, 0.1, 0.1, 0.1
    case, pos = stress_case(n, 20000)
    t = T.Tree([dict(name="chr1", anchors=case["anchors"], arcs=case["arcs"],
                     long_arcs=case["long_arcs"])], case["splits"])
    first, _ = t.chromosome(0)
    ids = list(range(first, first + n))
    p = T.level_params()
    runs = []
    for _ in range(3):
        t.set_positions(ids, pos)
        t0 = time.perf_counter()
        st = T.stress_refine(ctx, t, 0, p, ks, kq, ka, max_iter=iters, lr=lr)[0]
        runs.append(time.perf_counter() - t0)
    wall = min(runs)
    got = t.positions(ids)
    leg = {"workload": f"stress refinement of {n} anchors ({len(case['arcs'])} arcs, "
                       f"{len(case['long_arcs'])} long arcs), max {iters} iterations, step {lr}",
           "iterations": st["iterations"], "stop": st["stop"], "ms_per_call": wall * 1e3,
           "iterations_per_s": st["iterations"] / wall, "energy": st["total"]}
    if with_cpu:
        from oracle import refdriver
        if refdriver.available():
            params = {k: getattr(p, {"stretching": "heat_stretching"}.get(k, k))
                      for k in refdriver.PARAM_KEYS}
            ms, _, want, _ = refdriver.stress(case, params, ks, kq, ka, iters, lr, pos)
            leg["identical_to_reference"] = bool(np.array_equal(got.view(np.uint32),
                                                                want.view(np.uint32)))
            leg["cpu_reference"] = {
                "value": st["iterations"] / (ms * 1e-3), "unit": "iterations/s", "cores": 1,
                "kind": "reference", "ms_per_call": ms,
                "sample": "LooperSolver::stressRefineAnchors on the same chromosome and start, "
                          "one host thread (set-up included, as the call runs)"}
            leg["speedup_vs_reference"] = ms * 1e-3 / wall
    return leg


LEGS = {"heatmap_64": heatmap_leg, "heatmap_64_sweep": heatmap_sweep_leg,
        "smooth_19951": smooth_leg, "singleton_20M": singleton_leg, "stress_20k": stress_leg,
        "arcs_level": arcs_level_leg}
# legs runnable alone with --only (not part of `levels`)
ONLY = dict(LEGS, batched=lambda ctx, with_cpu=True: batched(ctx),
            matched_energy=lambda ctx, with_cpu=True: matched_energy(ctx),
            served=lambda ctx, with_cpu=True: served(ctx))


def levels(ctx, with_cpu=True, cpu_seconds=8.0):
    """The heatmap and smooth workloads, one solve each."""
    return {k: f(ctx, with_cpu, cpu_seconds) for k, f in LEGS.items()}


def batched(ctx, chains_per_problem=256, plain_subset=16):
    """Config 3: 512 problems in one pool vs one plain call per problem."""
    import numpy as np
    from pmmc_amd import schedule_preset, synth
    s = arcs_settings()
    probs, meta = synth.batch_arcs_problems(n_segments=32, n_seeds=16, seed=0)
    sched = schedule_preset(2, 1, s, 0.03)
    n = len(probs)
    seeds = [1 + i for i in range(n)]
    # one untimed call first (other seeds): it grows the library's device
    # arena to the pool's size, which a caller pays once per context
    ctx.anneal_batch(probs, sched, seeds=[100_000 + s for s in seeds], n_chains=chains_per_problem,
                     use_graph=True)
    # batched: launch shape chosen by the library's tuner (chains per
    # wavefront, blocks per step, steps per graph), timed with its probes
    shape = {}
    t0 = time.perf_counter()
    res = ctx.anneal_batch(probs, sched, seeds=seeds, n_chains=chains_per_problem, use_graph=True,
                           opts_out=shape)
    tb = time.perf_counter() - t0
    mv_b = sum(r[1]["moves"] for r in res)
    ac_b = sum(r[1]["accepted"] for r in res)
    # the same batch with one lane per chain at the chosen step shape
    t0 = time.perf_counter()
    res64 = ctx.anneal_batch(probs, sched, seeds=seeds, n_chains=chains_per_problem, use_graph=True,
                             blocks_per_step=shape["blocks_per_step"],
                             steps_per_graph=shape["steps_per_graph"], chains_per_wave=64)
    t64 = time.perf_counter() - t0
    # the same batch as one launch that runs every chain to its end
    # (blocks_per_step -1: no steps, graph or tuner; the workgroup dispatcher
    # starts ensembles beyond the resident ones as slots free up)
    t0 = time.perf_counter()
    res1 = ctx.anneal_batch(probs, sched, seeds=seeds, n_chains=chains_per_problem,
                            blocks_per_step=-1)
    t1 = time.perf_counter() - t0
    same_shape = all(np.array_equal(a[0], b[0]) and a[1]["energy"] == b[1]["energy"]
                     for a, b in zip(res, res64))
    same_shape &= all(np.array_equal(a[0], b[0]) and a[1]["energy"] == b[1]["energy"]
                      for a, b in zip(res, res1))
    # plain, one problem at a time, on an evenly spread subset
    idx = list(range(0, n, max(1, n // plain_subset)))[:plain_subset]
    t0 = time.perf_counter()
    same = True
    mv_p = ac_p = 0
    plain = {}
    for i in idx:
        pos, st = ctx.anneal(probs[i], sched, seed=1 + i, segment=i, n_chains=chains_per_problem)
        mv_p += st["moves"]
        ac_p += st["accepted"]
        plain[i] = (pos, st)
        same &= bool(np.array_equal(pos, res[i][0])) and st["energy"] == res[i][1]["energy"]
    tp = time.perf_counter() - t0
    # matched quality: final energy and the Pearson correlation of anchor
    # distances with the arc targets, batched vs plain, problem by problem
    from pmmc_amd import metrics

    def target_pearson(p, pos):
        ed = np.asarray(p.exp_dist, np.float32)
        a, b = np.nonzero(np.triu(ed > 1e-6, 1))
        d = np.linalg.norm(pos[a].astype(np.float64) - pos[b].astype(np.float64), axis=1)
        return metrics.pearson(d.astype(np.float32), ed[a, b])
    e_rel = max(abs(res[i][1]["energy"] - plain[i][1]["energy"]) / abs(plain[i][1]["energy"])
                for i in idx)
    p_diff = max(abs(target_pearson(probs[i], res[i][0]) - target_pearson(probs[i], plain[i][0]))
                 for i in idx)
    sub_moves_b = sum(res[i][1]["moves"] for i in idx)
    return {
        "problems": n, "chains_per_problem": chains_per_problem,
        "launch_shape": {k: shape[k] for k in ("chains_per_wave", "blocks_per_step",
                                                 "steps_per_graph", "tuned")},
        "one_lane_per_chain_moves_per_s": sum(r[1]["moves"] for r in res64) / t64,
        "one_launch_solves_per_s": n / t1,
        "results_identical_across_launch_shapes": same_shape,
        "anchors_min_max": [int(min(m[2] for m in meta)), int(max(m[2] for m in meta))],
        "batched_solves_per_s": n / tb, "batched_moves_per_s": mv_b / tb,
        "batched_accepted_moves_per_s": ac_b / tb,
        "plain_solves_per_s": len(idx) / tp, "plain_moves_per_s": mv_p / tp,
        "plain_accepted_moves_per_s": ac_p / tp,
        "solves_speedup": (n / tb) / (len(idx) / tp),
        "accepted_moves_speedup": (ac_b / tb) / (ac_p / tp),
        "plain_subset": len(idx), "subset_moves_equal": sub_moves_b == mv_p,
        "results_identical_alone_and_batched": same,
        "final_energy_max_rel_diff_vs_plain": float(e_rel),
        "pearson_to_targets_max_diff_vs_plain": float(p_diff),
        "final_energy_median": float(np.median([r[1]["energy"] for r in res])),
    }


def matched_energy(ctx, sample_chains=256):
    """The headline segment annealed to completion under the reference CPU
    annealer's own schedule (LooperSolver::MonteCarloArcs, the CPU preset:
    milestones every 50,000 moves, stop at the first one that fails), so that
    every GPU chain is the Markov chain the CPU baseline runs and a chain is
    one complete anneal, the CPU baseline's unit of work.  One plain call with
    the default chain count gives the throughput; `sample_chains`
    single-chain jobs give the distribution of one chain's final energy, to
    set against the CPU runs' (cpu_baseline.final_energy_median)."""
    import numpy as np
    from pmmc_amd import schedule_preset, settings, synth
    prob, _, step, _ = workload()
    sched = schedule_preset(2, 0, settings(**synth.CONFIG_SETTINGS[2]), step)
    t0 = time.perf_counter()
    _, st = ctx.anneal(prob, sched, seed=7, segment=0)
    wall = time.perf_counter() - t0
    res = ctx.anneal_batch([prob] * sample_chains, sched, seeds=[11] * sample_chains,
                           segments=[0] * sample_chains,
                           chain0s=[64 * k for k in range(sample_chains)], n_chains=1)
    e1 = np.array([r[1]["energy"] for r in res])
    return {"workload": "the headline segment (400 anchors) annealed to completion under "
                        "LooperSolver::MonteCarloArcs' schedule (default.ini [simulation_arcs])",
            "chains": st["n_chains"], "anneals_per_s": st["n_chains"] / wall,
            "solve_s": wall, "moves_per_chain_mean": st["moves"] / st["n_chains"],
            "best_chain_energy": st["energy"],
            "single_chain_sample": sample_chains,
            "single_chain_energy_median": float(np.median(e1)),
            "single_chain_energy_p10_p90": [float(np.percentile(e1, 10)),
                                            float(np.percentile(e1, 90))]}


def _alg_bytes_per_move(level, n, deg, acc_rate):
    """Algorithmic HBM bytes of one chain-move per kernel (the roofline
    accounting): heatmap 33-64 beads (register kernels): 12 B read + 12 B x
    acceptance written; other heatmaps: the N partners' and the moved bead's
    positions (12 B each) and the weight row (4 N B, shared by 64 chains);
    arcs: 12 B x (1 + deg) + 16 B x deg / 64 + 12 B x acceptance; smooth: the
    5-bead window (60 B) + the rest lengths (8 B / 64) + 12 B x acceptance."""
    if level == 1:
        if 32 < n <= 64:
            return 12.0 + 12.0 * acc_rate
        return 12.0 * (n + 1) + 4.0 * n / 64.0 + 12.0 * acc_rate
    if level == 2:
        return 12.0 * (1 + deg) + 16.0 * deg / 64.0 + 12.0 * acc_rate
    return 60.0 + 8.0 / 64.0 + 12.0 * acc_rate


def served(ctx, seconds=20.0, rate=200.0, seed=0, plain_sample=24, window=5.0, launch_ms=0.0,
           streams=0):
    """Config 4: Poisson arrivals of mixed jobs through the server, `seconds`
    of arrivals at `rate` per second; latency percentiles per `window` of
    arrival time show whether the server keeps up."""
    import numpy as np
    from pmmc_amd import schedule_preset, settings, synth
    rng = np.random.default_rng(seed)
    cap = 8192
    # GPU-kernel presets of default.ini; every solve capped at 8,192 moves per
    # chain (arcs and smooth: the preset's stop_condition_steps cap; heatmap:
    # max_moves) so a request's latency stays in the sub-second range
    hs = schedule_preset(1, 1, settings(steps=50000), 1.0)
    hs.max_moves = cap
    asched = schedule_preset(2, 1, settings(steps=cap, improvement=0.975, min_successes=100), 0.03)
    ss = schedule_preset(3, 1, settings(steps=cap, dt_temp=0.9995, improvement=0.99,
                                        min_successes=50), 0.3)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * seconds * 1.5)))
    arrivals = arrivals[arrivals < seconds]
    # a pool of distinct inputs per kind, drawn from by the requests (input
    # construction is not part of serving)
    pool = {"heatmap": [], "arcs": [], "smooth": []}
    for k in range(48):
        n = int(np.exp(rng.uniform(np.log(64), np.log(512))))
        pool["heatmap"].append((synth.heatmap_problem(n=n, seed=k)[0], hs, 1, n, 0.0))
        n = int(np.exp(rng.uniform(np.log(20), np.log(2000))))
        p = synth.arcs_problem(n=n, n_arcs=3 * n, seed=k)
        deg = float(np.count_nonzero(p.exp_dist > 1e-6)) / n
        pool["arcs"].append((p, asched, 2, n, deg))
        a = int(np.exp(rng.uniform(np.log(21), np.log(401))))
        sp = synth.smooth_problem(n_anchors=a, add=49, seed=k)
        pool["smooth"].append((sp, ss, 3, len(np.asarray(sp.pos).reshape(-1, 3)), 0.0))
    reqs = []
    for k, t in enumerate(arrivals):
        u = rng.random()
        ens = int(rng.integers(1, 17))
        kind = "heatmap" if u < 0.4 else ("arcs" if u < 0.8 else "smooth")
        reqs.append((float(t), pool[kind][int(rng.integers(0, 48))], ens))
    srv = ctx.server(hbm_budget=64 << 30, max_jobs=4096, blocks_per_step=0, launch_ms=launch_ms,
                     streams=streams)
    t0 = time.perf_counter()
    jobs = {}
    done_at = {}
    pending = set()
    nxt = 0
    steps = 0
    t_submit = t_step = 0.0
    while len(done_at) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt][0] <= now:
            t, (p, sc, _, _, _), ens = reqs[nxt]
            ts = time.perf_counter()
            jobs[nxt] = srv.submit_many(p, sc, seeds=[1 + e for e in range(ens)], segment=nxt,
                                        n_chains=64)
            t_submit += time.perf_counter() - ts
            pending.add(nxt)
            nxt += 1
        if not pending:
            if nxt < len(reqs):
                time.sleep(max(0.0, min(0.001, reqs[nxt][0] - now)))
            continue
        ts = time.perf_counter()
        n_done, _, _ = srv.step()
        t_step += time.perf_counter() - ts
        steps += 1
        if n_done:
            now = time.perf_counter() - t0
            for k in [k for k in pending if srv.done(jobs[k][-1]) and all(srv.done(j) for j in jobs[k])]:
                done_at[k] = now
                pending.discard(k)
    wall = time.perf_counter() - t0
    lat = np.array([done_at[k] - reqs[k][0] for k in range(len(reqs))]) * 1e3
    arr = np.array([r[0] for r in reqs])
    windows = []
    for w0 in np.arange(0.0, seconds, window):
        sel = (arr >= w0) & (arr < w0 + window)
        if sel.any():
            windows.append({"arrivals_s": [float(w0), float(w0 + window)], "requests": int(sel.sum()),
                            "p50_ms": float(np.percentile(lat[sel], 50)),
                            "p99_ms": float(np.percentile(lat[sel], 99))})
    solves = sum(len(v) for v in jobs.values())
    moves = acc_served = 0
    alg_bytes = 0.0
    for k, v in jobs.items():
        _, (_, _, level, n, deg), _ = reqs[k]
        for j in v:
            st = srv.result(j)[1]
            moves += st["moves"]
            acc_served += st["accepted"]
            ar = st["accepted"] / max(1, st["moves"])
            alg_bytes += st["moves"] * _alg_bytes_per_move(level, n, deg, ar)
    # the plain form on an evenly spread sample of the same requests, one
    # solve per call, results compared with the served ones
    sample = list(range(0, len(reqs), max(1, len(reqs) // plain_sample)))[:plain_sample]
    same = True
    n_plain = acc_plain = 0
    t1 = time.perf_counter()
    for k in sample:
        _, (p, sc, _, _, _), ens = reqs[k]
        for e in range(ens):
            pos, st = ctx.anneal(p, sc, seed=1 + e, segment=k, n_chains=64)
            spos, sst = srv.result(jobs[k][e])
            same &= bool(np.array_equal(pos, spos)) and st["energy"] == sst["energy"]
            n_plain += 1
            acc_plain += st["accepted"]
    tp = time.perf_counter() - t1
    srv.close()
    achieved = alg_bytes / wall / 1e9
    return {"requests": len(reqs), "arrival_rate_per_s": rate, "arrival_seconds": seconds,
            "solves": solves, "steps": steps, "wall_s": wall,
            "host_s": {"submit": t_submit, "step": t_step},
            "solves_per_s": solves / wall, "requests_per_s": len(reqs) / wall,
            "moves_per_s": moves / wall, "accepted_moves_per_s": acc_served / wall,
            "p50_latency_ms": float(np.percentile(lat, 50)),
            "p99_latency_ms": float(np.percentile(lat, 99)),
            "latency_per_window": windows,
            "hbm": {"achieved_GBps": achieved, "frac": achieved / HBM_PEAK_GBS,
                    "accounting": "algorithmic bytes per chain-move of each job's kernel x its "
                                  "moves, over the whole served run's wall time"},
            "move_cap_per_chain": cap,
            "schedules": "GPU-kernel presets of default.ini; every solve capped at 8,192 moves "
                         "per chain (arcs, smooth: the preset's stop_condition_steps cap; "
                         "heatmap: max_moves)",
            "plain_sample_requests": len(sample), "plain_solves_per_s": n_plain / tp,
            "plain_accepted_moves_per_s": acc_plain / tp,
            "solves_speedup_vs_plain": (solves / wall) / (n_plain / tp),
            "accepted_moves_speedup_vs_plain": (acc_served / wall) / (acc_plain / tp),
            "results_identical_alone_and_served": same,
            "mix": "40% heatmap 64-512 beads, 40% arcs 20-2000 anchors, 20% smooth 1k-20k beads; "
                   "ensemble 1-16 solves of 64 chains per request, inputs drawn from 48 of each "
                   "kind"}


def probe() -> int:
    """Sustained bandwidth of the move kernels' own access patterns
    (pmmc_probe_move_pattern), the ceiling their HBM figures are read
    against besides the nominal 8 TB/s: the arc move (7 random beads of
    768 B read, one written back by no lane, about half of the lanes, every
    lane) over a 1.25 GiB buffer like the 400-anchor workload's positions, at
    16 and 64 wavefronts per CU; the smooth move (one random 5-bead window,
    3,840 B; one lane writes one bead back in every other move, like a 1.4%
    acceptance rate) over 8 GiB, at 20 per CU; a streaming
    read; LDS read bandwidth."""
    from pmmc_amd import Context
    out = {"probe": "move patterns", "runs": []}
    with Context(0) as ctx:
        cus = ctx.cu_count
        for waves in (16 * cus, 64 * cus):
            for mode in (0, 1, 2):
                r = ctx.probe_move_pattern(5 << 28, waves, 1024, 7, 1, mode)
                out["runs"].append({"pattern": "arc move: 7 beads read", "write": ["none", "half lanes", "all lanes"][mode],
                                    "waves": waves, **r})
        # footprints that fit the 256 MB Infinity Cache
        for mib in (64, 256):
            r = ctx.probe_move_pattern(mib << 20, 16 * cus, 1024, 7, 1, 1)
            out["runs"].append({"pattern": f"arc move: 7 beads read, {mib} MiB buffer",
                                "write": "half lanes", "waves": 16 * cus, **r})
        for mode in (0, 3):
            r = ctx.probe_move_pattern(8 << 30, 20 * cus, 1024, 1, 5, mode)
            out["runs"].append({"pattern": "smooth move: 5-bead window read",
                                "write": {0: "none", 3: "one lane, every other move"}[mode],
                                "waves": 20 * cus, **r})
    print(json.dumps(out))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--chains", type=int, default=0, help="chains per solve (0 = library default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-levels", action="store_true")
    ap.add_argument("--no-matched", action="store_true")
    ap.add_argument("--probe", action="store_true",
                    help="only measure the move kernels' access-pattern bandwidth ceilings")
    ap.add_argument("--only", choices=sorted(ONLY),
                    help="run one level workload alone and print its JSON (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB three times (device to device), the known-size "
                         "transfer that scales the FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args(argv)
    if a.copy_calibration:
        return copy_calibration()
    if a.probe:
        return probe()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        return 2

    import numpy as np
    from pmmc_amd import Context, schedule_preset

    if a.only:
        with Context(0) as ctx:
            print(json.dumps({a.only: ONLY[a.only](ctx, with_cpu=not a.no_cpu)}))
        return 0

    prob, s, step, arcs = workload()
    sched = schedule_preset(2, 1, s, step)
    deg = 2.0 * len(arcs) / 400.0  # CSR entries per bead
    with Context(0) as ctx:
        for _ in range(a.warmup):
            ctx.anneal(prob, sched, seed=1, n_chains=a.chains)
        moves = acc = 0
        kern_ms = 0.0
        energies = []
        t0 = time.perf_counter()
        for k in range(a.steps):
            _, st = ctx.anneal(prob, sched, seed=100 + k, segment=k, n_chains=a.chains)
            moves += st["moves"]
            acc += st["accepted"]
            kern_ms += st["kernel_ms"]
            energies.append(st["energy"])
            chains = st["n_chains"]
        wall = time.perf_counter() - t0
        extra = {}
        if not a.no_batched:
            extra["batched"] = batched(ctx)
        if not a.no_served:
            extra["served"] = served(ctx)
        if not a.no_matched:
            extra["matched_energy"] = matched_energy(ctx)
        if not a.no_levels:
            extra["levels"] = levels(ctx, with_cpu=not a.no_cpu)
    value = moves / wall
    # algorithmic HBM bytes per chain-move: the moved bead's position and its
    # deg partners' positions (12 B each, read once for both evaluations),
    # the CSR row (deg x 16 B, one read shared by the 64 chains of an
    # ensemble) and the 12 B position write of an accepted move
    bytes_per_move = 12.0 * (1.0 + deg) + 16.0 * deg / 64.0 + 12.0 * acc / moves
    achieved = bytes_per_move * moves / (kern_ms * 1e-3) / 1e9
    tpm = measured_traffic_per_move()
    line = {
        "metric": METRIC, "value": value, "unit": "moves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain arc-level SA, one 2 Mb segment: 400 anchors, 1200 "
                               "PET-cluster arcs, GPU-kernel schedule of "
                               "default.ini [simulation_arcs]",
                   "n_anchors": 400, "n_arcs": int(len(arcs)), "chains_per_solve": chains,
                   "step_size": step},
        "accepted_moves_per_s": acc / wall,
        "solves_per_s": a.steps / wall,
        "kernel_moves_per_s": moves / (kern_ms * 1e-3),
        "final_energy_median": float(np.median(energies)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if tpm is None else tpm * moves / a.steps,
                     "bytes_per_move": bytes_per_move,
                     "accounting": "per chain-move: 12 B x (1 + deg) positions read + "
                                   "16 B x deg / 64 CSR row (shared by the 64 chains of an "
                                   "ensemble) + 12 B x acceptance rate written, deg = 2 x arcs / "
                                   "anchors; traffic = measured HBM bytes per launch (rocprofv3 "
                                   "FETCH_SIZE x2.0 + WRITE_SIZE x1.0, copy-kernel calibrated)"},
    }
    ceil = pattern_ceiling()
    if ceil:
        # the same random bead reads and write-backs with no arithmetic
        # (bench.py --probe, profiles/probe_move_patterns.txt)
        line["pattern_ceiling"] = {"moves_per_s": ceil, "frac": moves / (kern_ms * 1e-3) / ceil,
                                   "pattern": "7 random 768-byte bead reads + one bead written "
                                              "back by half of the lanes, per move"}
    line.update(extra)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(prob, s, step)
        cb = line["cpu_baseline"]
        if "runs_per_s" in cb and "matched_energy" in line:
            # complete anneals on both sides: the GPU's per second over the
            # host's, with each side's median final energy of one chain
            me = line["matched_energy"]
            me["cpu_anneals_per_s"] = cb["runs_per_s"]
            me["cpu_final_energy_median"] = cb["final_energy_median"]
            me["anneals_speedup_vs_cpu"] = me["anneals_per_s"] / cb["runs_per_s"]
            me["median_energy_rel_diff_vs_cpu"] = (me["single_chain_energy_median"] /
                                                   cb["final_energy_median"] - 1.0)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
