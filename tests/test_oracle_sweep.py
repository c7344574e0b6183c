"""Pins of the oracle's heatmap sweep (orc_heatmap_sweep), the restatement of
the reference GPU heatmap kernel's algorithm (ParallelMonteCarloHeatmap.cu:
217-363) as a synchronous sweep.  The reference kernel is CUDA with racing
writers, so it cannot be run here; what pins the restatement:

* an independent enumeration, written in Python from the sweep's definition
  (include/pmmc_hip.h, pmmc_heatmap_sweep), on tiny cases: n <= 4 beads,
  1-2 proposers, 1-4 proposals per iteration, several iterations, fixed
  beads, the sphere penalty, 2D moves;
* invariants the definition fixes: a bead moves only to a point whose local
  score against the snapshot is strictly lower, fixed beads never move, at
  T0 = 0 a proposer accepts no uphill proposal;
* the argmin's tie rule (lowest score, ties to the lowest proposer index,
  i.e. lowest lane, then lowest replica) on hand cases;
* the stop rule as a truth table over the energy trace: misses are counted
  and never reset (ParallelMonteCarloHeatmap.cu:343-345).
"""
import numpy as np
import pytest

from oracle import binding as O
from pmmc_amd.api import Heatmap


def philox_draw(seed, move, chain, segment, level=4, sub=0):
    key = (seed & 0xFFFFFFFF, seed >> 32)
    return O.philox([move, chain, segment, (level << 28) | sub], key)


def u01(x):
    return np.float32(int(x) >> 8) * np.float32(2.0 ** -24)


def sched(**kw):
    d = dict(T0=5.0, dt=0.99, improvement=0.99, small_energy=1e-4, tjs=50.0, tjc=20.0, step=0.75,
             step_decay=0.95, sphere_radius=0.0, inner=4, proposers=2, fails_threshold=3,
             max_iters=3, use2d=0)
    d.update(kw)
    return O.OrcSweep(d["T0"], d["dt"], d["improvement"], d["small_energy"], d["tjs"], d["tjc"],
                      d["step"], d["step_decay"], d["sphere_radius"], d["inner"], d["proposers"],
                      d["fails_threshold"], d["max_iters"], d["use2d"])


# ---------------- the independent enumeration ----------------
def _block(bnd, n, p):
    if bnd is None or len(bnd) == 0 or p < 0 or p > bnd[-1]:
        return 0, n - 1
    for i in range(len(bnd) - 1):
        if bnd[i] <= p < bnd[i + 1]:
            return bnd[i], bnd[i + 1] - 1
    return 0, n - 1


def _local(h, diag, bnd, thr, snap, b, x):
    """local score of bead b at x against the snapshot: sum over partners i of
    the block, outside the diagonal band, with h[i][b] >= thr, of
    (|x_i - x| / h[i][b] - 1)^2"""
    n = h.shape[0]
    st, en = _block(bnd, n, b)
    e = 0.0
    for i in range(st, en + 1):
        if i == b or abs(i - b) < diag or h[i, b] < thr:
            continue
        d = np.sqrt(sum((float(snap[i, k]) - float(x[k])) ** 2 for k in range(3)))
        c = d / float(h[i, b]) - 1.0
        e += c * c
    return e


def _sphere(x, r):
    if r <= 0:
        return 0.0
    d = np.sqrt(sum(float(v) ** 2 for v in x))
    return 0.0 if d <= r else (d - r) ** 2


def _full(h, diag, bnd, thr, pos):
    return sum(_local(h, diag, bnd, thr, pos, m, pos[m]) for m in range(h.shape[0]))


def py_sweep(h, diag, bnd, thr, fixed, pos0, S, seed, segment):
    """pmmc_heatmap_sweep's definition, step by step.  Returns (positions,
    energy trace, stop reason, iterations, accepted proposals, uphill
    proposals accepted)."""
    n = h.shape[0]
    pos = np.array(pos0, np.float32).copy()
    E = _full(h, diag, bnd, thr, pos)
    trace = [E]
    sprev = [E] * n                     # the first comparison is with the total
    T, ms, step = S.T0, E, np.float32(S.step)
    misses, it, acc, uphill = 0, 0, 0, 0
    if not (E == E) or E < 1e-6:
        return pos, trace, 4, 0, acc, uphill
    one, two = np.float32(1.0), np.float32(2.0)
    while True:
        if it >= S.max_iters:
            return pos, trace, 3, it, acc, uphill
        snap = pos.copy()
        for b in range(n):
            if fixed is not None and fixed[b]:
                continue
            finals = []
            for r in range(S.proposers):
                x = snap[b].copy()
                sp = sprev[b]
                for k in range(S.inner):
                    w = philox_draw(seed, it * S.inner + k, b * S.proposers + r, segment)
                    dx, dy, dz = ((two * u01(w[a]) - one) * step for a in range(3))
                    y = np.array([x[0] + dx, x[1] + dy, x[2] if S.use2d else x[2] + dz],
                                 np.float32)
                    sn = _local(h, diag, bnd, thr, snap, b, y) + _sphere(y, S.sphere_radius)
                    ok = sn <= sp
                    if not ok and T > 0:
                        ok = float(u01(w[3])) < S.tjs * np.exp(-S.tjc * (sn / sp) / T)
                        uphill += int(ok)
                    if ok:
                        x, sp = y, sn
                        acc += 1
                finals.append((sp, r, x))
            best_sc, _, best_x = min(finals, key=lambda t: (t[0], t[1]))
            lc = _local(h, diag, bnd, thr, snap, b, best_x)
            lg = _local(h, diag, bnd, thr, snap, b, snap[b])
            if lc < lg:
                pos[b] = best_x
            sprev[b] = best_sc
        T *= S.dt
        step = step * np.float32(S.step_decay)
        it += 1
        E = _full(h, diag, bnd, thr, pos)
        trace.append(E)
        if not np.isfinite(E):
            return pos, trace, 6, it, acc, uphill
        if E > S.improvement * ms:
            misses += 1                  # never reset
        if E < S.small_energy:
            return pos, trace, 2, it, acc, uphill
        if misses >= S.fails_threshold:
            return pos, trace, 1, it, acc, uphill
        ms = E


def _tiny(rng, n, fixed=False, bnd=None):
    h = rng.uniform(0.5, 3.0, (n, n)).astype(np.float32)
    h = ((h + h.T) / 2).astype(np.float32)
    np.fill_diagonal(h, 0)
    if n > 2:
        h[0, n - 1] = h[n - 1, 0] = 1e-4      # below the GPU skip threshold
    pos = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    fx = None
    if fixed:
        fx = np.zeros(n, np.uint8)
        fx[rng.integers(0, n)] = 1
    return Heatmap(pos=pos, dist=h, fixed=fx, boundaries=bnd, skip_below=1e-3)


CASES = [  # n, proposers, inner, iters, fixed, sphere, use2d, T0, boundaries
    (2, 1, 1, 1, False, 0.0, 0, 5.0, None),
    (2, 2, 4, 3, False, 0.0, 0, 5.0, None),
    (3, 2, 3, 3, True, 0.0, 0, 5.0, None),
    (3, 1, 4, 4, False, 1.5, 0, 2.0, None),
    (4, 2, 4, 3, True, 0.0, 1, 5.0, None),
    (4, 2, 2, 4, False, 0.0, 0, 0.0, None),
    (4, 2, 4, 3, False, 2.0, 0, 50.0, [0, 2, 4]),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_sweep_matches_enumeration(case):
    n, R, K, iters, fx, rad, u2, T0, bnd = CASES[case]
    rng = np.random.default_rng(100 + case)
    prob = _tiny(rng, n, fx, None if bnd is None else np.array(bnd, np.int32))
    S = sched(proposers=R, inner=K, max_iters=iters, sphere_radius=rad, use2d=u2, T0=T0,
              fails_threshold=10)
    P = O.Problem(prob)
    seed, seg = 1234 + case, 7
    got, st = O.sweep(P, S, seed, seg)
    want, trace, reason, it, acc, uphill = py_sweep(prob.dist, prob.diagonal, bnd, 1e-3,
                                                    prob.fixed, prob.pos, S, seed, seg)
    assert np.array_equal(got, want)
    assert st["milestones"] == it and st["stop_reason"] == reason
    assert st["accepted"] == acc
    nfree = n - (0 if prob.fixed is None else int(prob.fixed.sum()))
    assert st["moves"] == it * nfree * R * K
    assert np.allclose(st["e_trace"], trace, rtol=1e-13, atol=0)
    if T0 == 0.0:
        assert uphill == 0                        # at T0 = 0 nothing uphill passes


def test_sweep_enumeration_cases_move():
    # the cases above are not vacuous: most of them move beads
    moved = 0
    for case, (n, R, K, iters, fx, rad, u2, T0, bnd) in enumerate(CASES):
        rng = np.random.default_rng(100 + case)
        prob = _tiny(rng, n, fx, None if bnd is None else np.array(bnd, np.int32))
        S = sched(proposers=R, inner=K, max_iters=iters, sphere_radius=rad, use2d=u2, T0=T0,
                  fails_threshold=10)
        got, _ = O.sweep(O.Problem(prob), S, 1234 + case, 7)
        moved += not np.array_equal(got, prob.pos)
    assert moved >= len(CASES) - 2, moved


def _mid(seed=5, n=12):
    rng = np.random.default_rng(seed)
    walk = np.cumsum(rng.normal(size=(n, 3)), 0)
    d = np.linalg.norm(walk[:, None] - walk[None], axis=2).astype(np.float32)
    np.fill_diagonal(d, 0)
    pos = rng.uniform(0, 10, (n, 3)).astype(np.float32)
    fx = np.zeros(n, np.uint8)
    fx[[2, 7]] = 1
    return Heatmap(pos=pos, dist=d, fixed=fx, skip_below=1e-3)


@pytest.mark.parametrize("T0", [0.0, 5.0])
def test_sweep_moves_only_downhill_and_never_fixed(T0):
    prob = _mid()
    P = O.Problem(prob)
    h = prob.dist
    x = prob.pos.copy()
    moved_any = False
    for it in range(6):
        S = sched(T0=T0, proposers=4, inner=16, max_iters=1, fails_threshold=100)
        y, st = O.sweep(P, S, 77, it, pos=x)
        for b in range(prob.pos.shape[0]):
            if prob.fixed[b]:
                assert np.array_equal(y[b], x[b])
            elif not np.array_equal(y[b], x[b]):
                moved_any = True
                assert _local(h, 1, None, 1e-3, x, b, y[b]) < _local(h, 1, None, 1e-3, x, b, x[b])
        x = y
    assert moved_any


def test_sweep_argmin_ties():
    assert O.sweep_argmin([3.0, 1.0, 1.0, 2.0]) == 1
    assert O.sweep_argmin([1.0, 1.0]) == 0
    assert O.sweep_argmin([5.0]) == 0
    # proposer index = replica * 64 + lane: lane 3 of replica 1 (index 67)
    # loses to an equal score at lane 3 of replica 0 (index 3), and to an equal
    # score at lane 1 of replica 1 (index 65)
    s = np.full(128, 9.0)
    s[67] = s[3] = 0.5
    assert O.sweep_argmin(s) == 3
    s[3] = 0.6
    s[65] = 0.5
    assert O.sweep_argmin(s) == 65
    assert O.sweep_argmin([np.inf, 2.0, np.inf]) == 1


def _stop_iter(trace, impr, fails, small, reset):
    """iteration at which the stop rule fires on an energy trace (1-based),
    0 if it never does; misses never reset unless `reset`"""
    ms, misses = trace[0], 0
    for k in range(1, len(trace)):
        E = trace[k]
        if E > impr * ms:
            misses += 1
        elif reset:
            misses = 0
        if E < small or misses >= fails:
            return k
        ms = E
    return 0


def test_sweep_stop_rule_misses_never_reset():
    prob = _mid(11, n=10)
    P = O.Problem(prob)
    differs = 0
    for seed in range(12):
        S = sched(proposers=2, inner=8, max_iters=40, fails_threshold=3, improvement=0.999,
                  T0=20.0, dt=0.999)
        _, st = O.sweep(P, S, seed, 0)
        tr = st["e_trace"]
        it = st["milestones"]
        # replay on the full-length trace of an uncapped run of the same seed
        S2 = sched(proposers=2, inner=8, max_iters=40, fails_threshold=1000, improvement=0.999,
                   T0=20.0, dt=0.999)
        _, st2 = O.sweep(P, S2, seed, 0)
        full = st2["e_trace"]
        assert np.array_equal(full[:it + 1], tr)   # same run up to the stop
        k = _stop_iter(full, 0.999, 3, 1e-4, reset=False)
        if k:
            assert it == k and st["stop_reason"] == 1
        else:
            assert it == 40 and st["stop_reason"] == 3
        differs += _stop_iter(full, 0.999, 3, 1e-4, reset=True) != k
    # the traces tell the two rules apart (so the check above has teeth)
    assert differs > 0


def test_sweep_near_tie_count():
    prob = _mid(13, n=10)
    P = O.Problem(prob)
    S = sched(proposers=2, inner=8, max_iters=4, fails_threshold=100)
    x0, s0 = O.sweep(P, S, 9, 2)
    assert s0["near_ties"] == 0 and s0["first_near_tie"] == -1 and not s0["tie_beads"].any()
    x1, s1 = O.sweep(P, S, 9, 2, tie_rel=1e-6)
    assert np.array_equal(x0, x1)                 # counting never changes the run
    x2, s2 = O.sweep(P, S, 9, 2, tie_rel=1.0)     # every decision is a near tie
    assert np.array_equal(x0, x2)
    assert s2["first_near_tie"] == 0 and s2["near_ties"] == s2["milestones"]
    assert s2["tie_beads"][prob.fixed == 0].all() and not s2["tie_beads"][prob.fixed == 1].any()
