"""ctypes binding of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() (as the checker) and bench.py's
cpu_baseline leg (as the CPU baseline being timed).  Never by pmmc_amd/.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_MARKARCS = os.path.join(HERE, "_ref", "markarcs_ref")

LEVEL_HEATMAP, LEVEL_ARCS, LEVEL_SMOOTH, LEVEL_SWEEP = 1, 2, 3, 4

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)


class OrcSchedule(C.Structure):
    _fields_ = [
        ("T0", C.c_double), ("dt", C.c_double), ("improvement", C.c_double),
        ("small_energy", C.c_double), ("ratio_stop", C.c_double), ("guard_energy", C.c_double),
        ("tjs", C.c_float), ("tjc", C.c_float), ("step", C.c_float), ("step_decay", C.c_float),
        ("sphere_radius", C.c_float),
        ("block", C.c_int32), ("milestone_period", C.c_int32), ("min_successes", C.c_int32),
        ("fails_threshold", C.c_int32), ("stop_rule", C.c_int32), ("strict_accept", C.c_int32),
        ("positive_guard", C.c_int32), ("cool_per_block", C.c_int32), ("use2d", C.c_int32),
        ("resum", C.c_int32), ("max_moves", C.c_int64),
    ]


class OrcStats(C.Structure):
    _fields_ = [("energy", C.c_double), ("best_energy", C.c_double), ("T", C.c_double),
                ("moves", C.c_int64), ("accepted", C.c_int64), ("milestones", C.c_int32),
                ("stop_reason", C.c_int32), ("near_ties", C.c_int32), ("pad_", C.c_int32),
                ("first_near_tie", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class OrcHeatmap(C.Structure):
    _fields_ = [("n", C.c_int32), ("h", f32p), ("diag", C.c_int32), ("bnd", i32p),
                ("nb", C.c_int32), ("skip_below", C.c_float), ("fixed", u8p)]


class OrcArcs(C.Structure):
    _fields_ = [("n", C.c_int32), ("ed", f32p), ("nloops", C.c_int32), ("lpairs", i32p),
                ("lparams", f32p), ("ks", C.c_float), ("kq", C.c_float), ("fixed", u8p)]


class OrcSmooth(C.Structure):
    _fields_ = [("n", C.c_int32), ("dtn", f32p), ("ks", C.c_float), ("kq", C.c_float),
                ("ka", C.c_float), ("wd", C.c_float), ("wa", C.c_float), ("fixed", u8p)]


class OrcSweep(C.Structure):
    _fields_ = [("T0", C.c_double), ("dt", C.c_double), ("improvement", C.c_double),
                ("small_energy", C.c_double), ("tjs", C.c_float), ("tjc", C.c_float),
                ("step", C.c_float), ("step_decay", C.c_float), ("sphere_radius", C.c_float),
                ("inner", C.c_int32), ("proposers", C.c_int32), ("fails_threshold", C.c_int32),
                ("max_iters", C.c_int32), ("use2d", C.c_int32)]


class OrcStressStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("stop", C.c_int), ("lr", C.c_float),
                ("total", C.c_double), ("backbone", C.c_double), ("pairs", C.c_double),
                ("angular", C.c_double)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_sphere_penalty.restype = C.c_double
        L.orc_sphere_penalty.argtypes = [f32p, C.c_double]
        L.orc_heatmap_local.restype = C.c_double
        L.orc_heatmap_local.argtypes = [C.POINTER(OrcHeatmap), f32p, C.c_int]
        L.orc_heatmap_full.restype = C.c_double
        L.orc_heatmap_full.argtypes = [C.POINTER(OrcHeatmap), f32p]
        L.orc_heatmap_delta.restype = C.c_double
        L.orc_heatmap_delta.argtypes = [C.POINTER(OrcHeatmap), f32p, C.c_int, f32p, C.c_double]
        L.orc_arcs_total.restype = C.c_double
        L.orc_arcs_total.argtypes = [C.POINTER(OrcArcs), f32p, C.c_double]
        L.orc_arcs_repulsion.restype = C.c_double
        L.orc_arcs_repulsion.argtypes = [C.POINTER(OrcArcs), f32p]
        L.orc_arcs_local.restype = C.c_double
        L.orc_arcs_local.argtypes = [C.POINTER(OrcArcs), f32p, C.c_int, f32p, C.c_double]
        L.orc_smooth_full.restype = C.c_double
        L.orc_smooth_full.argtypes = [C.POINTER(OrcSmooth), f32p, C.c_double, C.c_int]
        L.orc_smooth_local.restype = C.c_double
        L.orc_smooth_local.argtypes = [C.POINTER(OrcSmooth), f32p, C.c_int, f32p, C.c_double]
        L.orc_anneal_chain.restype = C.c_int64
        L.orc_anneal_chain.argtypes = [C.c_int, C.c_void_p, C.POINTER(OrcSchedule), C.c_uint64,
                                       C.c_uint32, C.c_uint32, C.c_int, f32p,
                                       C.POINTER(OrcStats), u8p, C.c_int, C.c_double]
        L.orc_anneal_many.restype = C.c_double
        L.orc_anneal_many.argtypes = [C.c_int, C.c_void_p, C.POINTER(OrcSchedule), C.c_uint64,
                                      C.c_uint32, C.c_uint32, C.c_int, C.c_int, f32p, f32p,
                                      C.POINTER(OrcStats), C.c_int, C.c_int, C.c_double]
        L.orc_heatmap_sweep.restype = C.c_int
        L.orc_heatmap_sweep.argtypes = [C.POINTER(OrcHeatmap), C.POINTER(OrcSweep), C.c_uint64,
                                        C.c_uint32, f32p, C.POINTER(OrcStats), C.c_double, u8p,
                                        C.POINTER(C.c_double)]
        L.orc_sweep_argmin.restype = C.c_int
        L.orc_sweep_argmin.argtypes = [C.POINTER(C.c_double), C.c_int]
        L.orc_mark_arcs.argtypes = [i32p, i32p, C.c_int32, i32p, i32p, C.c_int32, i32p, i32p]
        L.orc_stress_gradient.argtypes = [C.c_int, f32p, f32p, C.c_int, i32p, i32p, f32p, C.c_float,
                                          C.c_float, C.c_float, f32p, C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_stress_refine.argtypes = [C.c_int, f32p, f32p, C.c_int, i32p, i32p, f32p, C.c_float,
                                        C.c_float, C.c_float, C.c_int, C.c_float,
                                        C.POINTER(OrcStressStats)]
        L.orc_stress_energy64.restype = C.c_double
        L.orc_stress_energy64.argtypes = [C.c_int, f32p, f32p, C.c_int, i32p, i32p, f32p,
                                          C.c_double, C.c_double, C.c_double]
        L.orc_acosf_sum.restype = C.c_uint64
        L.orc_acosf_sum.argtypes = [C.c_uint32, C.c_uint32, C.c_int]
        L.orc_group_arcs.restype = C.c_int
        L.orc_group_arcs.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, i32p, i32p, i32p, i32p,
                                     i32p]
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def philox(ctr, key):
    L = load()
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    L.orc_philox4x32_10(_p(c, u32p), _p(k, u32p), _p(o, u32p))
    return o


class Problem:
    """Oracle view of a pmmc_amd.api problem (Heatmap / Arcs / Smooth)."""

    def __init__(self, prob):
        self.level = prob.level
        self.pos0 = np.ascontiguousarray(prob.pos, np.float32).reshape(-1)
        self.n = self.pos0.size // 3
        fx = getattr(prob, "fixed", None)
        self.fixed = None if fx is None else np.ascontiguousarray(fx, np.uint8)
        if self.level == LEVEL_HEATMAP:
            self.h = np.ascontiguousarray(prob.dist, np.float32).reshape(-1)
            bd = prob.boundaries
            self.bnd = None if bd is None else np.ascontiguousarray(bd, np.int32)
            self.c = OrcHeatmap(self.n, _p(self.h, f32p), prob.diagonal, _p(self.bnd, i32p),
                                0 if self.bnd is None else self.bnd.size, prob.skip_below,
                                _p(self.fixed, u8p))
        elif self.level == LEVEL_ARCS:
            self.ed = np.ascontiguousarray(prob.exp_dist, np.float32).reshape(-1)
            lp = prob.loop_pairs
            self.lp = None if lp is None else np.ascontiguousarray(lp, np.int32).reshape(-1)
            self.lq = None if lp is None else np.ascontiguousarray(prob.loop_params, np.float32).reshape(-1)
            self.c = OrcArcs(self.n, _p(self.ed, f32p), 0 if self.lp is None else self.lp.size // 2,
                             _p(self.lp, i32p), _p(self.lq, f32p), prob.k_stretch, prob.k_squeeze,
                             _p(self.fixed, u8p))
        else:
            self.dtn = np.ascontiguousarray(prob.dist_to_next, np.float32)
            self.c = OrcSmooth(self.n, _p(self.dtn, f32p), prob.k_stretch, prob.k_squeeze,
                               prob.k_angular, prob.w_dist, prob.w_angle, _p(self.fixed, u8p))

    def ptr(self):
        return C.cast(C.pointer(self.c), C.c_void_p)

    # energies
    def total(self, pos=None, r=0.0):
        L = load()
        x = np.ascontiguousarray(self.pos0 if pos is None else pos, np.float32).reshape(-1)
        if self.level == LEVEL_HEATMAP:
            e = L.orc_heatmap_full(C.byref(self.c), _p(x, f32p))
            return e + sum(L.orc_sphere_penalty(_p(x[3 * i:3 * i + 3].copy(), f32p), r)
                           for i in range(self.n)) if r > 0 else e
        if self.level == LEVEL_ARCS:
            return L.orc_arcs_total(C.byref(self.c), _p(x, f32p), r)
        return L.orc_smooth_full(C.byref(self.c), _p(x, f32p), r, 1)

    def delta(self, pos, p, newp, r=0.0):
        L = load()
        x = np.ascontiguousarray(pos, np.float32).reshape(-1)
        q = np.ascontiguousarray(newp, np.float32)
        if self.level == LEVEL_HEATMAP:
            return L.orc_heatmap_delta(C.byref(self.c), _p(x, f32p), p, _p(q, f32p), r)
        if self.level == LEVEL_ARCS:
            f = L.orc_arcs_local
        else:
            f = L.orc_smooth_local
        return (f(C.byref(self.c), _p(x, f32p), p, _p(q, f32p), r)
                - f(C.byref(self.c), _p(x, f32p), p, None, r))

    def local_pair(self, pos, p, newp, r=0.0):
        """(local before, local after) magnitudes for the ΔE tolerance."""
        L = load()
        x = np.ascontiguousarray(pos, np.float32).reshape(-1)
        q = np.ascontiguousarray(newp, np.float32)
        if self.level == LEVEL_HEATMAP:
            y = x.copy()
            y[3 * p:3 * p + 3] = q
            a = L.orc_heatmap_local(C.byref(self.c), _p(x, f32p), p)
            b = L.orc_heatmap_local(C.byref(self.c), _p(y, f32p), p)
            return 2 * a, 2 * b
        f = L.orc_arcs_local if self.level == LEVEL_ARCS else L.orc_smooth_local
        return (f(C.byref(self.c), _p(x, f32p), p, None, r),
                f(C.byref(self.c), _p(x, f32p), p, _p(q, f32p), r))

    def pick_shift(self):
        """The stream rule: bead picks per chain (0) for heatmaps of 33..64
        beads, shared by the 64 chains of an ensemble (6) otherwise."""
        return 0 if self.level == LEVEL_HEATMAP and 32 < self.n <= 64 else 6

    def anneal(self, sched: OrcSchedule, seed, segment, chain, pick_shift=None, pos=None,
               trace=False, full_rescore=0, tie_rel=0.0):
        L = load()
        if pick_shift is None:
            pick_shift = self.pick_shift()
        x = (self.pos0 if pos is None else np.ascontiguousarray(pos, np.float32).reshape(-1)).copy()
        st = OrcStats()
        tr = np.zeros(int(min(sched.max_moves + sched.block, 1 << 26)), np.uint8) if trace else None
        L.orc_anneal_chain(self.level, self.ptr(), C.byref(sched), seed, segment, chain, pick_shift,
                           _p(x, f32p), C.byref(st), _p(tr, u8p), full_rescore, float(tie_rel))
        d = st.as_dict()
        if trace:
            d["trace"] = tr[:d["moves"]].copy()
        return x.reshape(-1, 3), d

    def anneal_many(self, sched, seed, segment, chain0, n_chains, n_threads, pick_shift=None,
                    full_rescore=0, tie_rel=0.0):
        L = load()
        if pick_shift is None:
            pick_shift = self.pick_shift()
        out = np.zeros(n_chains * self.n * 3, np.float32)
        stats = (OrcStats * n_chains)()
        secs = L.orc_anneal_many(self.level, self.ptr(), C.byref(sched), seed, segment, chain0,
                                 n_chains, pick_shift, _p(self.pos0, f32p), _p(out, f32p), stats,
                                 n_threads, full_rescore, float(tie_rel))
        return out.reshape(n_chains, self.n, 3), [s.as_dict() for s in stats], secs


def sweep(prob: "Problem", sched: OrcSweep, seed, segment, pos=None, tie_rel=0.0):
    """orc_heatmap_sweep: returns (positions [n,3], stats dict with
    'tie_beads' [n] bool and 'e_trace' [iterations + 1])."""
    L = load()
    x = (prob.pos0 if pos is None else np.ascontiguousarray(pos, np.float32).reshape(-1)).copy()
    st = OrcStats()
    tb = np.zeros(prob.n, np.uint8)
    tr = np.zeros(int(sched.max_iters) + 1, np.float64)
    it = L.orc_heatmap_sweep(C.byref(prob.c), C.byref(sched), seed, segment, _p(x, f32p),
                             C.byref(st), float(tie_rel), _p(tb, u8p),
                             tr.ctypes.data_as(C.POINTER(C.c_double)))
    d = st.as_dict()
    d["tie_beads"] = tb.astype(bool)
    d["e_trace"] = tr[:it + 1].copy()
    return x.reshape(-1, 3), d


def sweep_argmin(scores):
    L = load()
    a = np.ascontiguousarray(scores, np.float64)
    return L.orc_sweep_argmin(a.ctypes.data_as(C.POINTER(C.c_double)), a.size)


def schedule_from(s) -> OrcSchedule:
    """Copy a pmmc_amd Schedule (same field order) into an OrcSchedule."""
    o = OrcSchedule()
    for k, _ in OrcSchedule._fields_:
        setattr(o, k, getattr(s, k))
    return o


def mark_arcs(anc_s, anc_e, arc_s, arc_e):
    L = load()
    a_s, a_e = np.ascontiguousarray(anc_s, np.int32), np.ascontiguousarray(anc_e, np.int32)
    r_s, r_e = np.ascontiguousarray(arc_s, np.int32), np.ascontiguousarray(arc_e, np.int32)
    o_s, o_e = np.empty(r_s.size, np.int32), np.empty(r_s.size, np.int32)
    L.orc_mark_arcs(_p(a_s, i32p), _p(a_e, i32p), a_s.size, _p(r_s, i32p), _p(r_e, i32p),
                    r_s.size, _p(o_s, i32p), _p(o_e, i32p))
    return o_s, o_e


def group_arcs(chromosomes, use_unordered=True):
    L = load()
    anc_off, raw_off, anc, raw = [0], [0], [], []
    for a, r in chromosomes:
        a = np.asarray(a, np.int32).reshape(-1, 2)
        r = np.asarray(r, np.int32).reshape(-1, 4)
        anc.append(a)
        raw.append(r)
        anc_off.append(anc_off[-1] + len(a))
        raw_off.append(raw_off[-1] + len(r))
    A_ = np.concatenate(anc) if anc else np.zeros((0, 2), np.int32)
    R_ = np.ascontiguousarray(np.concatenate(raw) if raw else np.zeros((0, 4), np.int32))
    a_s, a_e = np.ascontiguousarray(A_[:, 0]), np.ascontiguousarray(A_[:, 1])
    ao, ro = np.asarray(anc_off, np.int32), np.asarray(raw_off, np.int32)
    out = np.zeros(max(1, 2 * raw_off[-1]) * 7, np.int32)
    oo = np.zeros(len(chromosomes) + 1, np.int32)
    nm = np.zeros(max(1, len(chromosomes)), np.int32)
    L.orc_group_arcs(1 if use_unordered else 0, len(chromosomes), _p(ao, i32p), _p(a_s, i32p),
                     _p(a_e, i32p), _p(ro, i32p), _p(R_.reshape(-1), i32p), _p(out, i32p),
                     _p(oo, i32p), _p(nm, i32p))
    out = out.reshape(-1, 7)
    return [out[oo[c]:oo[c + 1]].copy() for c in range(len(chromosomes))], nm[:len(chromosomes)]


def ref_mark_arcs(chromosomes):
    """Run the reference's own InteractionArcs::markArcs (oracle/_ref); None
    when that binary was not built (reference sources absent)."""
    import subprocess
    if not os.path.exists(REF_MARKARCS):
        return None
    lines = [str(len(chromosomes))]
    for a, r in chromosomes:
        a = np.asarray(a).reshape(-1, 2)
        r = np.asarray(r).reshape(-1, 4)
        lines.append(f"{len(a)} {len(r)}")
        lines += [f"{x} {y}" for x, y in a.tolist()]
        lines += [" ".join(map(str, row)) for row in r.tolist()]
    p = subprocess.run([REF_MARKARCS], input="\n".join(lines) + "\n", stdout=subprocess.PIPE,
                       stderr=subprocess.DEVNULL, text=True, check=True)
    vals = p.stdout.split()
    out, k = [], 0
    for _ in chromosomes:
        m = int(vals[k])
        k += 1
        arr = np.array(vals[k:k + 7 * m], np.int64).reshape(m, 7).astype(np.int32)
        k += 7 * m
        out.append(arr)
    return out


def _stress_args(pos, bb, pairs):
    x = np.ascontiguousarray(pos, np.float32).reshape(-1, 3).copy()
    n = len(x)
    b = np.ascontiguousarray(bb, np.float32) if n > 1 else np.zeros(1, np.float32)
    pr = np.asarray(pairs, np.float64).reshape(-1, 3)
    pi = np.ascontiguousarray(pr[:, 0], np.int32)
    pj = np.ascontiguousarray(pr[:, 1], np.int32)
    pt = np.ascontiguousarray(pr[:, 2], np.float32)
    return x, n, b, pi, pj, pt


def stress_refine(pos, bb, pairs, ks, kq, ka, max_iter, lr, targets=None):
    """orc_stress_refine: stressRefineAnchors' descent; pairs [(i, j, t)]
    (targets, when given, replace t as exact float32 values).  Returns
    (positions [n,3], stats dict)."""
    x, n, b, pi, pj, pt = _stress_args(pos, bb, pairs)
    if targets is not None:
        pt = np.ascontiguousarray(targets, np.float32)
    st = OrcStressStats()
    load().orc_stress_refine(n, _p(x, f32p), _p(b, f32p), len(pi), _p(pi, i32p), _p(pj, i32p),
                             _p(pt, f32p), ks, kq, ka, int(max_iter), lr, C.byref(st))
    return x, {k: getattr(st, k) for k, _ in st._fields_}


def stress_gradient(pos, bb, pairs, ks, kq, ka):
    """orc_stress_gradient: (gradient [n,3], e_back, e_hic, e_ang)."""
    x, n, b, pi, pj, pt = _stress_args(pos, bb, pairs)
    G = np.zeros((n, 3), np.float32)
    e = [C.c_double(), C.c_double(), C.c_double()]
    load().orc_stress_gradient(n, _p(x, f32p), _p(b, f32p), len(pi), _p(pi, i32p), _p(pj, i32p),
                               _p(pt, f32p), ks, kq, ka, _p(G, f32p), *[C.byref(v) for v in e])
    return G, e[0].value, e[1].value, e[2].value


def stress_energy64(pos, bb, pairs, ks, kq, ka):
    x, n, b, pi, pj, pt = _stress_args(pos, bb, pairs)
    return load().orc_stress_energy64(n, _p(x, f32p), _p(b, f32p), len(pi), _p(pi, i32p),
                                      _p(pj, i32p), _p(pt, f32p), ks, kq, ka)


def acosf_sum(first_bits, count, threads=8):
    """orc_acosf_sum: libm acosf checksum over a range of float bit patterns."""
    return int(load().orc_acosf_sum(int(first_bits), int(count), int(threads)))
