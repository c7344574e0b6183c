"""Benchmark: plain arc-level simulated annealing of one 2 Mb segment
(400 anchors, 1,200 PET-cluster arcs) on one MI355X.

One step = one plain solve (pmmc_arcs_anneal: all chains of the problem run to
their stop rule, the best chain is returned).  Inputs are uploaded inside the
step, as the drop-in does; `value` is counted moves per second over the timed
steps (whole job).  The roofline figure uses the annealing kernel's own
duration, measured with HIP events on the library's stream.
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


def genomic_step(anchors, noise=0.01, power=0.75, scale=0.5, base=1.0):
    """noise_size of LooperSolver::reconstructClusterArcsDistances
    (LooperSolver.hpp:4339-4354): mean genomicLengthToDistance of consecutive
    anchors times noise_arcs ([distance] and [main] of default.ini)."""
    import numpy as np
    g = anchors.mean(axis=1)
    d = base + scale * np.power(np.diff(g) / 1000.0, power)
    return float(d.sum() / len(anchors) * noise)


def workload():
    from pmmc_amd import schedule_preset, settings, synth
    anchors, arcs, ed = synth.arcs_inputs(n=400, n_arcs=1200, seed=0)
    prob = synth.arcs_problem(n=400, n_arcs=1200, seed=0)
    step = genomic_step(anchors)
    # [simulation_arcs] and [cuda] of auxiliary/configs/default.ini
    s = settings(max_temp=5.0, dt_temp=0.9999, improvement=0.975, tjs=50.0, tjc=20.0,
                 min_successes=100, steps=50000, milestone_fails=3)
    return prob, s, step, arcs


def cpu_baseline(prob, s, step, budget_s=15.0):
    """The reference's CPU annealer (LooperSolver::MonteCarloArcs, dense O(n)
    local scans) restated in the oracle, one chain per host thread."""
    from oracle import binding as O
    from pmmc_amd import schedule_preset
    sched = O.schedule_from(schedule_preset(2, 0, s, step))
    P = O.Problem(prob)
    threads = max(1, min(16, os.cpu_count() or 1))
    # calibrate on one thread per core, then size the sample to the budget
    moves_per_chain = 50000
    sched.max_moves = moves_per_chain
    sched.block = 50000
    sched.milestone_period = 50000
    _, st, secs = P.anneal_many(sched, 1, 0, 0, threads, threads)
    per_chain = secs  # one chain per thread ran concurrently
    rounds = max(1, int(budget_s / max(per_chain, 1e-3)))
    chains = threads * rounds
    _, stats, secs = P.anneal_many(sched, 1, 0, 0, chains, threads)
    moves = sum(x["moves"] for x in stats)
    return {"value": moves / secs, "unit": "moves/s", "cores": threads, "kind": "port",
            "sample": f"{chains} chains x {moves_per_chain} moves of the CPU arc annealer "
                      f"(LooperSolver::MonteCarloArcs schedule, dense O(n) local scans), "
                      f"{secs:.1f} s on {threads} host threads"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--chains", type=int, default=0, help="chains per solve (0 = 64 per CU)")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        return 2

    import numpy as np
    from pmmc_amd import Context, schedule_preset

    prob, s, step, arcs = workload()
    sched = schedule_preset(2, 1, s, step)
    deg = 2.0 * len(arcs) / 400.0  # CSR entries per bead
    bytes_per_move = 2.0 * deg * (12 + 4 + 4)
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
    value = moves / wall
    achieved = bytes_per_move * moves / (kern_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC, "value": value, "unit": "moves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain arc-level SA, one 2 Mb segment: 400 anchors, 1200 "
                               "PET-cluster arcs (configs[1]), GPU-kernel schedule of "
                               "default.ini [simulation_arcs]",
                   "n_anchors": 400, "n_arcs": int(len(arcs)), "chains_per_solve": chains,
                   "step_size": step},
        "accepted_moves_per_s": acc / wall,
        "solves_per_s": a.steps / wall,
        "kernel_moves_per_s": moves / (kern_ms * 1e-3),
        "final_energy_median": float(np.median(energies)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_move": bytes_per_move,
                     "accounting": "CSR variant: 2 evaluations x deg x (12 B position + 4 B "
                                   "target + 4 B index) per proposal"},
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(prob, s, step)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
