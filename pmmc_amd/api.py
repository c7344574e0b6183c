"""numpy-facing wrappers over the C ABI (include/pmmc_hip.h).

Every call goes through libpmmc_hip.so; arrays are converted to contiguous
float32 / int32 / uint8 and kept alive for the duration of the call.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A


class PmmcError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{A.STATUS.get(status, status)}: {msg}")
        self.status = status


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def settings(max_temp=5.0, dt_temp=0.9999, improvement=0.99, tjs=50.0, tjc=20.0,
             min_successes=10, steps=50000, milestone_fails=3, sphere_radius=0.0, use2d=0):
    """pmmc_level_settings from [simulation_*] / [cuda] values."""
    return A.LevelSettings(max_temp, dt_temp, improvement, tjs, tjc, min_successes, steps,
                           milestone_fails, sphere_radius, use2d)


def schedule_preset(level: int, preset: int, s: A.LevelSettings, step_size: float) -> A.Schedule:
    lib = A.load()
    out = A.Schedule()
    st = lib.pmmc_schedule_preset(level, preset, C.byref(s), step_size, C.byref(out))
    if st != 0:
        raise PmmcError(st, "bad level or preset")
    return out


def pick_shift(level: int, n: int) -> int:
    """pmmc_pick_shift: chains c and c' share bead picks iff c >> s == c' >> s."""
    return int(A.load().pmmc_pick_shift(int(level), int(n)))


@dataclass
class Heatmap:
    pos: np.ndarray
    dist: np.ndarray
    fixed: np.ndarray | None = None
    diagonal: int = 1
    boundaries: np.ndarray | None = None
    skip_below: float = 1e-6
    _keep: list = field(default_factory=list, repr=False)

    level = A.LEVEL_HEATMAP

    def c(self):
        pos, dist = _f32(self.pos).reshape(-1), _f32(self.dist).reshape(-1)
        fx = None if self.fixed is None else _u8(self.fixed)
        bd = None if self.boundaries is None else _i32(self.boundaries)
        self._keep = [pos, dist, fx, bd]
        n = pos.size // 3
        return A.HeatmapProblem(n, _ptr(pos, A.f32p), _ptr(fx, A.u8p), _ptr(dist, A.f32p),
                                self.diagonal, _ptr(bd, A.i32p), 0 if bd is None else bd.size,
                                self.skip_below)


@dataclass
class Arcs:
    pos: np.ndarray
    exp_dist: np.ndarray
    fixed: np.ndarray | None = None
    loop_pairs: np.ndarray | None = None
    loop_params: np.ndarray | None = None
    k_stretch: float = 1.0
    k_squeeze: float = 1.0
    _keep: list = field(default_factory=list, repr=False)

    level = A.LEVEL_ARCS

    def c(self):
        pos, ed = _f32(self.pos).reshape(-1), _f32(self.exp_dist).reshape(-1)
        fx = None if self.fixed is None else _u8(self.fixed)
        lp = None if self.loop_pairs is None else _i32(self.loop_pairs).reshape(-1)
        lq = None if self.loop_params is None else _f32(self.loop_params).reshape(-1)
        self._keep = [pos, ed, fx, lp, lq]
        nl = 0 if lp is None else lp.size // 2
        return A.ArcsProblem(pos.size // 3, _ptr(pos, A.f32p), _ptr(fx, A.u8p), _ptr(ed, A.f32p),
                             nl, _ptr(lp, A.i32p), _ptr(lq, A.f32p), self.k_stretch,
                             self.k_squeeze)


@dataclass
class Smooth:
    pos: np.ndarray
    dist_to_next: np.ndarray
    fixed: np.ndarray | None = None
    k_stretch: float = 0.1
    k_squeeze: float = 0.1
    k_angular: float = 0.1
    w_dist: float = 1.0
    w_angle: float = 1.0
    _keep: list = field(default_factory=list, repr=False)

    level = A.LEVEL_SMOOTH

    def c(self):
        pos, dt = _f32(self.pos).reshape(-1), _f32(self.dist_to_next)
        fx = None if self.fixed is None else _u8(self.fixed)
        self._keep = [pos, dt, fx]
        return A.SmoothProblem(pos.size // 3, _ptr(pos, A.f32p), _ptr(fx, A.u8p),
                               _ptr(dt, A.f32p), self.k_stretch, self.k_squeeze, self.k_angular,
                               self.w_dist, self.w_angle)


class Context:
    """One pmmc_ctx on one device."""

    def __init__(self, device: int = 0):
        self.lib = A.load()
        h = C.c_void_p()
        st = self.lib.pmmc_ctx_create(device, C.byref(h))
        if st != 0:
            raise PmmcError(st, f"pmmc_ctx_create(device={device}) failed")
        self.h = h

    def close(self):
        if self.h:
            self.lib.pmmc_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            msg = self.lib.pmmc_last_error(self.h).decode() or self.lib.pmmc_host_last_error().decode()
            raise PmmcError(st, msg)

    def spline_positions(self, genomic_pos, xyz, first_end: int, last_start: int, pos_start: int,
                         resolution_bp: int, n_out: int | None = None, pos_stop: int | None = None):
        """pmmc_spline_positions: Catmull-Rom points at pos_start + k*resolution_bp
        (n_out points, or the reference's count for [pos_start, pos_stop])."""
        gp = np.ascontiguousarray(genomic_pos, np.int32)
        x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        if n_out is None:
            n_out = self.lib.pmmc_spline_count(int(pos_start), int(pos_stop), int(resolution_bp))
        out = np.empty((int(n_out), 3), np.float32)
        self._check(self.lib.pmmc_spline_positions(
            self.h, len(gp), gp.ctypes.data_as(A.i32p), x.ctypes.data_as(A.f32p), int(first_end),
            int(last_start), int(pos_start), int(resolution_bp), int(n_out),
            out.ctypes.data_as(A.f32p)))
        return out

    def stress_refine(self, problems, max_iter: int, lr: float = 0.05, in_order_sums=False):
        """pmmc_stress_refine over a list of dicts {pos [n,3], backbone [n-1],
        pairs [(i, j, target), ...] (or arrays pair_i, pair_j, targets),
        k_stretch, k_squeeze, k_angular}, one launch; returns
        [(positions, stats dict), ...]."""
        keep, cps, outs = [], [], []
        for q in problems:
            pos = np.ascontiguousarray(q["pos"], np.float32).reshape(-1, 3)
            n = len(pos)
            bb = np.ascontiguousarray(q.get("backbone", np.full(max(n - 1, 1), -1.0)), np.float32)
            if "pair_i" in q:      # arrays: pair_i, pair_j, targets
                pi = np.ascontiguousarray(q["pair_i"], np.int32)
                pj = np.ascontiguousarray(q["pair_j"], np.int32)
                pt = np.ascontiguousarray(q["targets"], np.float32)
            else:
                pr = np.asarray(q.get("pairs", []), np.float64).reshape(-1, 3)
                pi = np.ascontiguousarray(pr[:, 0], np.int32)
                pj = np.ascontiguousarray(pr[:, 1], np.int32)
                pt = np.ascontiguousarray(q.get("targets", pr[:, 2]), np.float32)
            out = pos.copy()
            keep += [pos, bb, pi, pj, pt]
            outs.append(out)
            cps.append(A.StressProblem(n, pos.ctypes.data_as(A.f32p), bb.ctypes.data_as(A.f32p),
                                       len(pi), pi.ctypes.data_as(A.i32p), pj.ctypes.data_as(A.i32p),
                                       pt.ctypes.data_as(A.f32p), float(q.get("k_stretch", 0.1)),
                                       float(q.get("k_squeeze", 0.1)),
                                       float(q.get("k_angular", 0.1))))
        arr = (A.StressProblem * max(1, len(cps)))(*cps)
        optr = (A.f32p * max(1, len(outs)))(*[o.ctypes.data_as(A.f32p) for o in outs])
        st = (A.StressStats * max(1, len(cps)))()
        o = A.StressOpts(int(max_iter), float(lr), 1 if in_order_sums else 0)
        self._check(self.lib.pmmc_stress_refine(self.h, len(cps), arr, C.byref(o), optr, st))
        return [(outs[i], st[i].as_dict()) for i in range(len(cps))]

    def probe_acosf(self, first_bits: int, count: int) -> int:
        """pmmc_probe_acosf: the device arc cosine's checksum over a range of
        float bit patterns."""
        out = C.c_uint64()
        self._check(self.lib.pmmc_probe_acosf(self.h, int(first_bits), int(count), C.byref(out)))
        return int(out.value)

    def singleton_heatmap(self, chrs, clusters, records, region=(0, 0), centromeres=(),
                          long_arcs=(), long_scale=10.0, long_power=2.0, diag=1):
        """pmmc_singleton_heatmap: contact heatmap of a level's clusters from
        singleton records.  chrs: chromosome names in order; clusters: per
        chromosome the level's (start, end) in level order; records: either
        (chr1, start1, end1, chr2, start2, end2, score) tuples with chromosome
        names, or a dict of int32 arrays chr1/start1/end1/chr2/start2/end2/score
        with chromosome indices (-1: not processed); centromeres: (chr, start,
        end), the first per chromosome counts; long_arcs: (chr, start, end,
        score).  Returns (heat [N, N] float32, counts dict)."""
        idx = {c: k for k, c in enumerate(chrs)}
        off = np.zeros(len(chrs) + 1, np.int32)
        for k, cl in enumerate(clusters):
            off[k + 1] = off[k] + len(cl)
        cs = np.array([c[0] for cl in clusters for c in cl], np.int32)
        ce = np.array([c[1] for cl in clusters for c in cl], np.int32)
        if isinstance(records, dict):
            rec = {k: np.ascontiguousarray(records[k], np.int32)
                   for k in ("chr1", "start1", "end1", "chr2", "start2", "end2", "score")}
        else:
            rl = list(records)
            rec = {"chr1": np.array([idx.get(r[0], -1) for r in rl], np.int32),
                   "start1": np.array([r[1] for r in rl], np.int32),
                   "end1": np.array([r[2] for r in rl], np.int32),
                   "chr2": np.array([idx.get(r[3], -1) for r in rl], np.int32),
                   "start2": np.array([r[4] for r in rl], np.int32),
                   "end2": np.array([r[5] for r in rl], np.int32),
                   "score": np.array([r[6] for r in rl], np.int32)}
        cent = None
        if centromeres:
            cent = np.zeros(3 * len(chrs), np.int32)
            for c, a, b in centromeres:
                k = idx.get(c, -1)
                if k >= 0 and cent[3 * k] == 0:
                    cent[3 * k:3 * k + 3] = (1, a, b)
        la = list(long_arcs)
        lchr = np.array([idx.get(a[0], -1) for a in la], np.int32)
        lst = np.array([a[1] for a in la], np.int32)
        len_ = np.array([a[2] for a in la], np.int32)
        lsc = np.array([a[3] for a in la], np.int32)
        n = int(off[-1])
        inp = A.SingletonInput(
            len(chrs), _ptr(off, A.i32p), _ptr(cs, A.i32p), _ptr(ce, A.i32p), len(rec["chr1"]),
            *(_ptr(rec[k], A.i32p) for k in ("chr1", "start1", "end1", "chr2", "start2", "end2", "score")),
            int(region[0]), int(region[1]), _ptr(cent, A.i32p), len(la), _ptr(lchr, A.i32p),
            _ptr(lst, A.i32p), _ptr(len_, A.i32p), _ptr(lsc, A.i32p), float(long_scale),
            float(long_power), int(diag))
        heat = np.empty((n, n), np.float32)
        cnt = A.SingletonCounts()
        self._check(self.lib.pmmc_singleton_heatmap(self.h, C.byref(inp), _ptr(heat, A.f32p),
                                                    C.byref(cnt)))
        return heat, cnt.as_dict()

    def probe_move_pattern(self, bytes_: int, waves: int, moves_per_wave: int, reads: int,
                           chunk_beads: int, write_mode: int) -> dict:
        """pmmc_probe_move_pattern: sustained bandwidth of the move kernels'
        access pattern (random 768-byte bead reads, one bead written back),
        plus a streaming read and the LDS read bandwidth."""
        r = A.ProbeResult()
        self._check(self.lib.pmmc_probe_move_pattern(self.h, int(bytes_), int(waves),
                                                      int(moves_per_wave), int(reads),
                                                      int(chunk_beads), int(write_mode),
                                                      C.byref(r)))
        return r.as_dict()

    @property
    def cu_count(self) -> int:
        return self.lib.pmmc_ctx_cu_count(self.h)

    def set_gpu_seed(self, seed: int):
        self.lib.pmmc_set_gpu_seed(self.h, seed)

    def default_chains(self, level: int, n: int, symmetric: bool = True) -> int:
        """pmmc_default_chains: chains a call with n_chains = 0 runs."""
        return int(self.lib.pmmc_default_chains(self.h, int(level), int(n), int(bool(symmetric))))

    def reserve(self, nbytes: int):
        """pmmc_ctx_reserve: grow the context's device arena to nbytes."""
        self._check(self.lib.pmmc_ctx_reserve(self.h, C.c_size_t(int(nbytes))))

    # ---- plain ----
    def anneal(self, prob, sched: A.Schedule, seed=1, segment=0, chain0=0, n_chains=0):
        """Anneal one problem; returns (positions [n,3], stats dict)."""
        cp = prob.c()
        n = cp.n
        out = _f32(prob.pos).reshape(n, 3).copy()
        stats = A.Stats()
        run = A.Run(seed, segment, chain0, n_chains)
        fn = {A.LEVEL_HEATMAP: self.lib.pmmc_heatmap_anneal, A.LEVEL_ARCS: self.lib.pmmc_arcs_anneal,
              A.LEVEL_SMOOTH: self.lib.pmmc_smooth_anneal}[prob.level]
        self._check(fn(self.h, C.byref(cp), C.byref(sched), C.byref(run),
                       out.ctypes.data_as(A.f32p), C.byref(stats)))
        return out, stats.as_dict()

    def anneal_noalloc(self, prob, sched: A.Schedule, seed=1, segment=0, chain0=0, n_chains=0):
        """pmmc_anneal_noalloc: as anneal(), with no device allocation (the
        library's static 64 MiB scratch per device)."""
        cp = prob.c()
        n = cp.n
        out = _f32(prob.pos).reshape(n, 3).copy()
        stats = A.Stats()
        run = A.Run(seed, segment, chain0, n_chains)
        self._check(self.lib.pmmc_anneal_noalloc(self.h, prob.level, C.cast(C.pointer(cp), C.c_void_p),
                                                 C.byref(sched), C.byref(run),
                                                 out.ctypes.data_as(A.f32p), C.byref(stats)))
        return out, stats.as_dict()

    def delta_energy(self, prob, beads, disp, sphere_radius=0.0):
        cp = prob.c()
        b = _i32(beads)
        d = _f32(disp).reshape(-1)
        out = np.empty(b.size, np.float32)
        self._check(self.lib.pmmc_delta_energy(self.h, prob.level, C.cast(C.pointer(cp), C.c_void_p),
                                               sphere_radius, b.size, _ptr(b, A.i32p),
                                               _ptr(d, A.f32p), _ptr(out, A.f32p)))
        return out

    def heatmap_sweep(self, prob: Heatmap, sched: A.SweepSchedule, seed=1, segment=0):
        cp = prob.c()
        out = _f32(prob.pos).reshape(-1, 3).copy()
        stats = A.Stats()
        self._check(self.lib.pmmc_heatmap_sweep(self.h, C.byref(cp), C.byref(sched), seed, segment,
                                                out.ctypes.data_as(A.f32p), C.byref(stats)))
        return out, stats.as_dict()

    # ---- arcs marking ----
    def mark_arcs(self, anc_s, anc_e, arc_s, arc_e):
        a_s, a_e, r_s, r_e = _i32(anc_s), _i32(anc_e), _i32(arc_s), _i32(arc_e)
        o_s = np.empty(r_s.size, np.int32)
        o_e = np.empty(r_s.size, np.int32)
        self._check(self.lib.pmmc_mark_arcs(self.h, _ptr(a_s, A.i32p), _ptr(a_e, A.i32p), a_s.size,
                                            _ptr(r_s, A.i32p), _ptr(r_e, A.i32p), r_s.size,
                                            _ptr(o_s, A.i32p), _ptr(o_e, A.i32p)))
        return o_s, o_e

    def group_arcs(self, chromosomes):
        """chromosomes: list of (anchors [A,2], raw arcs [R,4]) ->
        list of arrays [K,7] (start, end, gstart, gend, score, eff, factor),
        list of non-matching counts."""
        anc_off, raw_off = [0], [0]
        anc, raw = [], []
        for a, r in chromosomes:
            a = np.asarray(a, np.int32).reshape(-1, 2)
            r = np.asarray(r, np.int32).reshape(-1, 4)
            anc.append(a)
            raw.append(r)
            anc_off.append(anc_off[-1] + len(a))
            raw_off.append(raw_off[-1] + len(r))
        A_ = np.concatenate(anc) if anc else np.zeros((0, 2), np.int32)
        R_ = np.concatenate(raw) if raw else np.zeros((0, 4), np.int32)
        a_s, a_e = _i32(A_[:, 0]), _i32(A_[:, 1])
        rr = _i32(R_).reshape(-1)
        ao, ro = _i32(anc_off), _i32(raw_off)
        out = (A.Arc * max(1, 2 * raw_off[-1]))()
        out_off = np.zeros(len(chromosomes) + 1, np.int32)
        nonm = np.zeros(max(1, len(chromosomes)), np.int32)
        self._check(self.lib.pmmc_group_arcs(self.h, len(chromosomes), _ptr(ao, A.i32p),
                                             _ptr(a_s, A.i32p), _ptr(a_e, A.i32p), _ptr(ro, A.i32p),
                                             _ptr(rr, A.i32p), out, _ptr(out_off, A.i32p),
                                             _ptr(nonm, A.i32p)))
        flat = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_int32)),
                                     shape=(max(1, 2 * raw_off[-1]) * 7,)).reshape(-1, 7)
        res = [flat[out_off[c]:out_off[c + 1]].copy() for c in range(len(chromosomes))]
        return res, nonm[:len(chromosomes)].copy()

    # ---- served ----
    def server(self, hbm_budget=8 << 30, max_jobs=256, blocks_per_step=4, launch_ms=0.0,
               streams=0):
        return Server(self, hbm_budget, max_jobs, blocks_per_step, launch_ms, streams)

    # ---- batched ----
    def anneal_batch(self, problems, schedules, seeds, segments=None, n_chains=0,
                     blocks_per_step=0, steps_per_graph=0, use_graph=True, chain0s=None,
                     chains_per_wave=0, opts_out=None):
        """Run many problems concurrently; returns list of (pos, stats).
        Job i uses seed seeds[i], segment segments[i] (default i) and chain ids
        chain0s[i] (default 0) onwards.  Launch-shape fields left 0 are
        autotuned; opts_out (a dict, optional) receives the values used."""
        k = len(problems)
        jobs = (A.Job * k)()
        keep, outs, stats = {}, [], []
        for i, p in enumerate(problems):
            # one problem struct per problem object: the library prepares a
            # problem that several jobs name once
            if id(p) not in keep:
                keep[id(p)] = (p, p.c())
            cp = keep[id(p)][1]
            o = _f32(p.pos).reshape(-1, 3).copy()
            s = A.Stats()
            outs.append(o)
            stats.append(s)
            seg = i if segments is None else segments[i]
            nc = n_chains[i] if isinstance(n_chains, (list, tuple, np.ndarray)) else n_chains
            sch = schedules[i] if isinstance(schedules, (list, tuple)) else schedules
            jobs[i] = A.Job(p.level, C.cast(C.pointer(cp), C.c_void_p), sch,
                            A.Run(int(seeds[i]), int(seg), 0 if chain0s is None else int(chain0s[i]),
                                  int(nc)),
                            o.ctypes.data_as(A.f32p), C.pointer(s))
        opts = A.BatchOpts(blocks_per_step, steps_per_graph, 1 if use_graph else 0,
                           chains_per_wave, 0)
        self._check(self.lib.pmmc_anneal_batch(self.h, jobs, k, C.byref(opts)))
        if opts_out is not None:
            opts_out.update(opts.as_dict())
        return [(outs[i], stats[i].as_dict()) for i in range(k)]


class Server:
    """Continuous-batching server (pmmc_server_*): submit jobs, step, collect."""

    def __init__(self, ctx: Context, hbm_budget: int, max_jobs: int, blocks_per_step: int,
                 launch_ms: float = 0.0, streams: int = 0):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        opts = A.ServerOpts(hbm_budget, max_jobs, blocks_per_step, launch_ms, streams)
        ctx._check(self.lib.pmmc_server_create(ctx.h, C.byref(opts), C.byref(h)))
        self.h = h
        self.jobs = {}

    def submit(self, prob, sched: A.Schedule, seed=1, segment=0, n_chains=0):
        cp = prob.c()
        out = _f32(prob.pos).reshape(-1, 3).copy()
        st = A.Stats()
        job = A.Job(prob.level, C.cast(C.pointer(cp), C.c_void_p), sched,
                    A.Run(seed, segment, 0, n_chains), out.ctypes.data_as(A.f32p), C.pointer(st))
        jid = C.c_uint64()
        self.ctx._check(self.lib.pmmc_server_submit(self.h, C.byref(job), C.byref(jid)))
        self.jobs[jid.value] = (prob, cp, out, st)
        return jid.value

    def submit_many(self, prob, sched: A.Schedule, seeds, segment=0, n_chains=0):
        """One job per seed on the same problem (pmmc_server_submit_many: the
        problem is prepared once); returns the job ids."""
        cp = prob.c()
        k = len(seeds)
        jobs = (A.Job * k)()
        outs = []
        for i, sd in enumerate(seeds):
            out = _f32(prob.pos).reshape(-1, 3).copy()
            st = A.Stats()
            jobs[i] = A.Job(prob.level, C.cast(C.pointer(cp), C.c_void_p), sched,
                            A.Run(int(sd), segment, 0, n_chains), out.ctypes.data_as(A.f32p),
                            C.pointer(st))
            outs.append((out, st))
        ids = (C.c_uint64 * k)()
        self.ctx._check(self.lib.pmmc_server_submit_many(self.h, jobs, k, ids))
        for i in range(k):
            self.jobs[ids[i]] = (prob, cp, outs[i][0], outs[i][1])
        return list(ids)

    def step(self):
        d, r, q = C.c_int32(), C.c_int32(), C.c_int32()
        self.ctx._check(self.lib.pmmc_server_step(self.h, C.byref(d), C.byref(r), C.byref(q)))
        return d.value, r.value, q.value

    def done(self, jid) -> bool:
        return bool(self.lib.pmmc_server_done(self.h, jid))

    def result(self, jid):
        _, _, out, st = self.jobs[jid]
        return out, st.as_dict()

    def close(self):
        if self.h:
            self.lib.pmmc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
