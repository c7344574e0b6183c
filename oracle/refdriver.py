"""Runs the reference's own CPU annealers -- test and CPU-baseline infrastructure only.

oracle/_ref/levels_ref is the reference's LooperSolver compiled in place from
its sources by oracle/Makefile (target `ref`; only where the reference
sources exist, the binary then travels with the tree).  This module formats
its `anneal` / `time_anneal` / `singleton_clusters` / `stress` input and
reads its output.  Used by tests/golden/make_anneal_fixtures.py (fixtures) and
bench.py's CPU legs (the reference's CPU annealing, and its singleton
heatmap construction, timed on the host's cores).
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "_ref", "levels_ref")
HEADER = "0 " * 12  # count/heatmap/genomic-length parameters (unused by these modes)


def available() -> bool:
    return os.path.exists(REF) and os.access(REF, os.X_OK)


def hx(a) -> str:
    return " ".join(float(v).hex() for v in np.asarray(a, np.float32).reshape(-1))


def anneal_input(level: int, prob, step: float, settings: dict, seed: int) -> str:
    """Input of the driver's anneal modes for one problem (pmmc_amd.api
    Heatmap / Arcs / Smooth) and [simulation_*] settings."""
    s = settings
    n = len(np.asarray(prob.pos).reshape(-1, 3))
    parts = [f"{level} {seed}", float(np.float32(step)).hex(),
             f"{s['max_temp']} {s['dt_temp']} {s['tjs']} {s['tjc']} {s['improvement']} "
             f"{s['min_successes']} {s['steps']} 0"]
    if level == 1:
        fixed = np.zeros(n, np.uint8) if prob.fixed is None else prob.fixed
        parts += [f"{n} {prob.diagonal}", hx(prob.pos), " ".join(str(int(v)) for v in fixed),
                  hx(prob.dist)]
    elif level == 2:
        parts += [f"{n} {prob.k_stretch} {prob.k_squeeze}", hx(prob.pos), hx(prob.exp_dist)]
    else:
        fixed = np.zeros(n, np.uint8) if prob.fixed is None else prob.fixed
        parts += [f"{n} {prob.k_stretch} {prob.k_squeeze} {prob.k_angular} {prob.w_dist} "
                  f"{prob.w_angle}", hx(prob.pos), " ".join(str(int(v)) for v in fixed),
                  hx(prob.dist_to_next)]
    return "\n".join(parts) + "\n"


def run(text: str, tag: str) -> list[str]:
    r = subprocess.run([REF], input=text, capture_output=True, text=True, check=True)
    lines = r.stdout.splitlines()
    k = [i for i, line in enumerate(lines) if line.startswith(tag)][0]
    return lines[k:]


def anneal(level: int, prob, step: float, settings: dict, seed: int):
    """One run of the reference annealer: (returned score, moves, positions)."""
    out = run(f"anneal {HEADER}\n" + anneal_input(level, prob, step, settings, seed), "@@anneal")
    _, score, moves = out[0].split()
    n = len(np.asarray(prob.pos).reshape(-1, 3))
    pos = np.array([[float.fromhex(t) for t in line.split()] for line in out[1:1 + n]], np.float32)
    return float.fromhex(score), int(moves), pos


def time_anneal(level: int, prob, step: float, settings: dict, seconds: float, threads: int):
    """`threads` processes each annealing repeatedly (seeds 1, 2, ...) for about
    `seconds` of annealing time.  Returns (moves, annealing seconds per
    process (max), runs, the score each run returned): whole-host moves/s =
    moves / seconds."""
    text = f"time_anneal {HEADER}\n{seconds}\n"
    procs = []
    for t in range(threads):
        inp = text + anneal_input(level, prob, step, settings, 1 + 1000 * t)
        p = subprocess.Popen([REF], stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                             stderr=subprocess.DEVNULL, text=True)
        p.stdin.write(inp)
        p.stdin.close()
        procs.append(p)
    moves, secs, runs, scores = 0, 0.0, 0, []
    for p in procs:
        out = p.stdout.read()
        if p.wait() != 0:
            raise RuntimeError("reference annealer failed")
        lines = out.splitlines()
        line = [x for x in lines if x.startswith("@@time")][0]
        _, r, m, s = line.split()
        runs += int(r)
        moves += int(m)
        secs = max(secs, float(s))
        scores += [float.fromhex(t) for t in [x for x in lines if x.startswith("@@scores")][0].split()[1:]]
    return moves, secs, runs, scores


def singleton_time(chrs, clusters, rec, diag=1, long_scale=10.0, long_power=2.0) -> tuple[float, int]:
    """LooperSolver::createSingletonHeatmap run once by the reference on the
    level's clusters and records (dict of int32 arrays chr1/start1/end1/chr2/
    start2/end2/score, chromosome indices, -1: a chromosome not processed),
    single-threaded as the reference runs it, reading them from a singleton
    text file as it does.  Returns (seconds inside the call, records)."""
    names = list(chrs) + ["chrUn"]
    parts = [f"singleton_clusters {HEADER}\n{len(chrs)}\n"]
    for c, cl in zip(chrs, clusters):
        parts.append(f"{c} {len(cl)}\n" + "".join(f"{a} {b}\n" for a, b in cl))
    parts.append(f"{diag}\n{float(np.float32(long_scale))} {float(np.float32(long_power))}\n0 0\n0\n0\n")
    n = len(rec["chr1"])
    cols = [np.asarray(rec[k]) for k in ("chr1", "start1", "end1", "chr2", "start2", "end2", "score")]
    c1 = np.array(names, dtype=object)[cols[0]]
    c2 = np.array(names, dtype=object)[cols[3]]
    lines = [f"{a} {b} {c} {d} {e} {f} {g}" for a, b, c, d, e, f, g in
             zip(c1, cols[1].tolist(), cols[2].tolist(), c2, cols[4].tolist(), cols[5].tolist(),
                 cols[6].tolist())]
    parts.append(f"{n}\n" + "\n".join(lines) + "\n")
    out = run("".join(parts), "@@singleton_time")
    _, secs, cnt = out[0].split()
    return float(secs), int(cnt)


PARAM_KEYS = ["count_a", "count_scale", "count_shift", "count_base", "heat_scale", "heat_power",
              "heat_scale_inter", "heat_power_inter", "stretching", "glen_base", "glen_scale",
              "glen_power"]


def stress(case: dict, params: dict, ks: float, kq: float, ka: float, iters: int, lr: float, pos):
    """LooperSolver::stressRefineAnchors run once by the reference on one
    chromosome: case {anchors [(start, end)], arcs [(a, b, count)] with local
    anchor indices, long_arcs [(start, end, count)], splits [(chr, start,
    end)]} named chr1, the anchors' start positions pos [n, 3].  Returns
    (milliseconds inside the call, sorted anchor order, positions [n, 3],
    the function's progress lines)."""
    f32 = lambda x: repr(float(np.float32(x)))  # noqa: E731
    c = case
    text = "stress " + " ".join(f32(params[k]) for k in PARAM_KEYS) + "\n"
    text += f"1\nchr1 {len(c['anchors'])} {len(c['arcs'])}\n"
    text += "".join(f"{s} {e} N\n" for s, e in c["anchors"])
    text += "".join(f"{a} {b} {sc}\n" for a, b, sc in c["arcs"])
    text += f"{len(c['splits'])}\n" + "".join(f"{n} {s} {e}\n" for n, s, e in c["splits"])
    text += f"chr1\n{f32(ks)} {f32(kq)} {f32(ka)}\n{iters} {f32(lr)}\n{len(c['long_arcs'])}\n"
    text += "".join(f"{s} {e} {sc}\n" for s, e, sc in c["long_arcs"])
    text += " ".join(float(v).hex() for v in np.asarray(pos, np.float32).reshape(-1)) + "\n"
    r = subprocess.run([REF], input=text, capture_output=True, text=True, check=True)
    lines = r.stdout.splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.startswith("@@stress"))
    n = len(c["anchors"])
    ms = float(lines[k].split()[2])
    order = [int(x) for x in lines[k + 1].split()]
    out = np.array([[float.fromhex(v) for v in ln.split()] for ln in lines[k + 2:k + 2 + n]],
                   np.float32)
    log = [ln for ln in lines[:k] if ln.startswith("[stressRefine]")]
    return ms, order, out, log
