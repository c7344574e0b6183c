"""numpy front end of include/pmmc_levels.h: the cluster tree, densification,
expected distances and the distance heatmap (host code of libpmmc_hip.so).

    from pmmc_amd.tree import Tree, level_params
    t = Tree([{"name": "chr1", "anchors": [(s, e), ...], "arcs": [(a, b, count), ...]}],
             splits=[("chr1", 0, 2_000_000), ...])
    ibs = t.level(0, 2)                  # interaction blocks of chromosome 0
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Sequence

import numpy as np

from . import _abi as A

_i32p = A.i32p
_f32p = A.f32p


def _lib():
    return A.load()


def _check(st, what):
    if st != 0:
        msg = _lib().pmmc_host_last_error().decode()
        raise RuntimeError(f"{what}: {A.STATUS.get(st, st)}: {msg}")


def level_params(**overrides) -> A.LevelParams:
    """Reference defaults (Settings.hpp), with keyword overrides."""
    p = A.LevelParams()
    _lib().pmmc_level_params_default(C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def freq_to_distance(count: int, p: A.LevelParams) -> np.float32:
    return np.float32(_lib().pmmc_freq_to_distance(int(count), C.byref(p)))


def genomic_length_to_distance(length: int, p: A.LevelParams) -> float:
    return _lib().pmmc_genomic_length_to_distance(int(length), C.byref(p))


def distance_heatmap(freq, diagonal_size: int, inter: bool, p: A.LevelParams):
    """Returns (distance heatmap, average used for the clamp)."""
    f = np.ascontiguousarray(freq, np.float32)
    n = f.shape[0]
    out = np.empty((n, n), np.float32)
    avg = C.c_float()
    _check(_lib().pmmc_distance_heatmap(f.ctypes.data_as(_f32p), n, int(diagonal_size), int(inter),
                                        C.byref(p), out.ctypes.data_as(_f32p), C.byref(avg)),
           "pmmc_distance_heatmap")
    return out, np.float32(avg.value)


def normalize_heatmap(heat):
    """pmmc_normalize_heatmap (LooperSolver::normalizeHeatmap)."""
    h = np.ascontiguousarray(heat, np.float32)
    out = np.empty_like(h)
    _check(_lib().pmmc_normalize_heatmap(h.ctypes.data_as(_f32p), h.shape[0],
                                         out.ctypes.data_as(_f32p)), "pmmc_normalize_heatmap")
    return out


def normalize_heatmap_diagonal_total(heat, val=1.0):
    """pmmc_normalize_heatmap_diagonal_total (normalizeHeatmapDiagonalTotal)."""
    h = np.array(heat, np.float32)
    _check(_lib().pmmc_normalize_heatmap_diagonal_total(h.ctypes.data_as(_f32p), h.shape[0],
                                                        float(val)),
           "pmmc_normalize_heatmap_diagonal_total")
    return h


def normalize_heatmap_inter(heat, sizes, scale):
    """pmmc_normalize_heatmap_inter (normalizeHeatmapInter); sizes: bins per
    chromosome."""
    h = np.array(heat, np.float32)
    sz = np.ascontiguousarray(sizes, np.int32)
    _check(_lib().pmmc_normalize_heatmap_inter(h.ctypes.data_as(_f32p), h.shape[0], len(sz),
                                               sz.ctypes.data_as(C.POINTER(C.c_int32)),
                                               float(scale)),
           "pmmc_normalize_heatmap_inter")
    return h


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, np.int64).astype(np.int32))


class Tree:
    """Owns a pmmc_tree.  chrs: [{name, anchors [(start, end)], arcs
    [(a, b, count)] (chromosome-local anchor indices), optional orient,
    long_arcs [(start_bp, end_bp, count)]}]; splits: [(chr, start, end)]."""

    def __init__(self, chrs: Sequence[dict], splits: Iterable = ()):
        self._keep = []
        cs = (A.Chromosome * max(1, len(chrs)))()
        for k, c in enumerate(chrs):
            anc = np.asarray(c["anchors"], np.int64).reshape(-1, 2)
            arcs = np.asarray(c.get("arcs", []), np.int64).reshape(-1, 3)
            longs = np.asarray(c.get("long_arcs", []), np.int64).reshape(-1, 3)
            arrs = [_i32(anc[:, 0]), _i32(anc[:, 1]), _i32(arcs[:, 0]), _i32(arcs[:, 1]),
                    _i32(arcs[:, 2]), _i32(longs[:, 0]), _i32(longs[:, 1]), _i32(longs[:, 2])]
            name = c["name"].encode()
            orient = "".join(c["orient"]).encode() if c.get("orient") else None
            self._keep += arrs + [name, orient]
            p = [a.ctypes.data_as(_i32p) for a in arrs]
            cs[k] = A.Chromosome(name, len(anc), p[0], p[1], orient, len(arcs), p[2], p[3], p[4],
                                 len(longs), p[5], p[6], p[7])
        splits = list(splits)
        rs = (A.SegmentRegion * max(1, len(splits)))()
        for k, (chr_name, s, e) in enumerate(splits):
            b = chr_name.encode()
            self._keep.append(b)
            rs[k] = A.SegmentRegion(b, int(s), int(e))
        h = C.c_void_p()
        _check(_lib().pmmc_tree_create(len(chrs), cs, len(splits), rs, C.byref(h)),
               "pmmc_tree_create")
        self.handle = h
        self._keep = []

    @classmethod
    def from_hcm(cls, path: str, fmt: int = 2) -> "Tree":
        """pmmc_tree_read_hcm: fmt 0 previous layout, 1 new, 2 the
        reference's detection rule."""
        t = cls.__new__(cls)
        t._keep = []
        h = C.c_void_p()
        _check(_lib().pmmc_tree_read_hcm(str(path).encode(), int(fmt), C.byref(h)), "pmmc_tree_read_hcm")
        t.handle = h
        return t

    def set_factors(self, names):
        arr = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
        _check(_lib().pmmc_tree_set_factors(self.handle, len(names), arr), "pmmc_tree_set_factors")

    def set_arc_info(self, chr_index, arcs):
        """arcs: [(start, end, genomic_start, genomic_end, score, eff_score,
        factor)] in the chromosome's arc order."""
        arr = (A.Arc * max(1, len(arcs)))(*[A.Arc(*map(int, a)) for a in arcs])
        _check(_lib().pmmc_tree_set_arc_info(self.handle, int(chr_index), len(arcs), arr),
               "pmmc_tree_set_arc_info")

    def set_annotation(self, ids, types):
        i = _i32(ids)
        t = _i32(types)
        _check(_lib().pmmc_tree_set_annotation(self.handle, len(i), i.ctypes.data_as(_i32p),
                                               t.ctypes.data_as(_i32p)), "pmmc_tree_set_annotation")

    def write_hcm(self, path, fmt: int = 1):
        _check(_lib().pmmc_tree_write_hcm(self.handle, str(path).encode(), int(fmt)),
               "pmmc_tree_write_hcm")

    def write_pdb(self, path, level: int = -1):
        _check(_lib().pmmc_tree_write_pdb(self.handle, int(level), str(path).encode()),
               "pmmc_tree_write_pdb")

    def write_cif(self, path, level: int = -1, data_name: str = "cudaMMC"):
        _check(_lib().pmmc_tree_write_cif(self.handle, int(level), data_name.encode(),
                                          str(path).encode()), "pmmc_tree_write_cif")

    def close(self):
        if self.handle:
            _lib().pmmc_tree_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return _lib().pmmc_tree_size(self.handle)

    @property
    def n_chromosomes(self):
        return _lib().pmmc_tree_n_chromosomes(self.handle)

    def clusters(self, first=0, count=None):
        count = len(self) - first if count is None else count
        out = (A.Cluster * max(1, count))()
        _check(_lib().pmmc_tree_get(self.handle, first, count, out), "pmmc_tree_get")
        return [out[i] for i in range(count)]

    def _list(self, fn, *args):
        n = C.c_int32()
        _check(fn(self.handle, *args, None, 0, C.byref(n)), fn.__name__)
        out = np.empty(max(1, n.value), np.int32)
        _check(fn(self.handle, *args, out.ctypes.data_as(_i32p), n.value, C.byref(n)), fn.__name__)
        return out[:n.value].tolist()

    def children(self, cluster):
        return self._list(_lib().pmmc_tree_children, int(cluster))

    def level(self, chr_index, level):
        return self._list(_lib().pmmc_tree_level, int(chr_index), int(level))

    def chromosome(self, chr_index):
        f, r = C.c_int32(), C.c_int32()
        _check(_lib().pmmc_tree_chromosome(self.handle, int(chr_index), C.byref(f), C.byref(r)),
               "pmmc_tree_chromosome")
        return f.value, r.value

    def arcs(self, chr_index):
        n = C.c_int32()
        _check(_lib().pmmc_tree_arcs(self.handle, int(chr_index), None, 0, C.byref(n)), "pmmc_tree_arcs")
        out = np.empty(max(1, 2 * n.value), np.int32)
        _check(_lib().pmmc_tree_arcs(self.handle, int(chr_index), out.ctypes.data_as(_i32p), n.value,
                                     C.byref(n)), "pmmc_tree_arcs")
        return out[:2 * n.value].reshape(-1, 2).tolist()

    def set_positions(self, ids, xyz):
        ids = _i32(ids)
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        _check(_lib().pmmc_tree_set_positions(self.handle, len(ids), ids.ctypes.data_as(_i32p),
                                              xyz.ctypes.data_as(_f32p)), "pmmc_tree_set_positions")

    def positions(self, ids):
        ids = _i32(ids)
        out = np.empty((len(ids), 3), np.float32)
        _check(_lib().pmmc_tree_get_positions(self.handle, len(ids), ids.ctypes.data_as(_i32p),
                                              out.ctypes.data_as(_f32p)), "pmmc_tree_get_positions")
        return out

    def set_fixed(self, ids, fixed):
        ids, fx = _i32(ids), _i32(fixed)
        _check(_lib().pmmc_tree_set_fixed(self.handle, len(ids), ids.ctypes.data_as(_i32p),
                                          fx.ctypes.data_as(_i32p)), "pmmc_tree_set_fixed")

    def densify(self, cluster, loop_density, fix=True):
        n = C.c_int32()
        _check(_lib().pmmc_tree_densify(self.handle, int(cluster), int(loop_density), int(bool(fix)),
                                        C.byref(n)), "pmmc_tree_densify")
        return n.value

    def set_dist_to_next(self, cluster, p):
        a, s = C.c_float(), C.c_float()
        _check(_lib().pmmc_tree_set_dist_to_next(self.handle, int(cluster), C.byref(p), C.byref(a),
                                                 C.byref(s)), "pmmc_tree_set_dist_to_next")
        return np.float32(a.value), np.float32(s.value)

    def expected_distances(self, chr_index, cluster, p, anchor_heatmap=None):
        n = len(self.children(cluster))
        out = np.empty((n, n), np.float32)
        h = None
        if anchor_heatmap is not None:
            h = np.ascontiguousarray(anchor_heatmap, np.float32)
            if h.shape != (n, n):
                raise ValueError(f"anchor heatmap must be {n}x{n}")
        _check(_lib().pmmc_tree_expected_distances(
            self.handle, int(chr_index), int(cluster), C.byref(p),
            h.ctypes.data_as(_f32p) if h is not None else None, out.ctypes.data_as(_f32p)),
            "pmmc_tree_expected_distances")
        return out


# ------------------------------------------------------------------ level drivers
def arcs_level_opts(**overrides) -> A.ArcsLevelOpts:
    """Reference defaults (Settings.hpp), with keyword overrides; `arcs` and
    `smooth` take pmmc_amd.api.settings(...) objects."""
    o = A.ArcsLevelOpts()
    _lib().pmmc_arcs_level_opts_default(C.byref(o))
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def heatmap_level_opts(**overrides) -> A.HeatmapLevelOpts:
    o = A.HeatmapLevelOpts()
    _lib().pmmc_heatmap_level_opts_default(C.byref(o))
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def stress_refine(ctx, tree: Tree, chr_index: int, p: A.LevelParams, k_stretch=0.1,
                  k_squeeze=0.1, k_angular=0.1, max_iter=0, lr=0.05, in_order_sums=False):
    """pmmc_tree_stress_refine on a pmmc_amd.api.Context: the anchors of
    chromosome chr_index (-1: all, one launch) refined in place; returns the
    stats dict of each chromosome processed."""
    n = tree.n_chromosomes if chr_index < 0 else 1
    st = (A.StressStats * max(1, n))()
    o = A.StressOpts(int(max_iter), float(lr), 1 if in_order_sums else 0)
    ctx._check(_lib().pmmc_tree_stress_refine(ctx.h, tree.handle, int(chr_index), C.byref(p),
                                              float(k_stretch), float(k_squeeze), float(k_angular),
                                              C.byref(o), st))
    return [st[i].as_dict() for i in range(n)]


def _report(r: A.LevelReport) -> dict:
    return {k: getattr(r, k) for k, _ in A.LevelReport._fields_}


def arcs_level(ctx, tree: Tree, chr_index: int, p: A.LevelParams, o: A.ArcsLevelOpts,
               seed: int = 1, segment: int = 0) -> dict:
    """pmmc_arcs_level on a pmmc_amd.api.Context; positions land in the tree."""
    rep = A.LevelReport()
    ctx._check(_lib().pmmc_arcs_level(ctx.h, tree.handle, int(chr_index), C.byref(p), C.byref(o),
                                      int(seed), int(segment), C.byref(rep)))
    return _report(rep)


def heatmap_level(ctx, tree: Tree, level: int, dist, diagonal: int, p: A.LevelParams,
                  o: A.HeatmapLevelOpts, seed: int = 1, segment: int = 0) -> dict:
    d = np.ascontiguousarray(dist, np.float32)
    rep = A.LevelReport()
    ctx._check(_lib().pmmc_heatmap_level(ctx.h, tree.handle, int(level), d.ctypes.data_as(_f32p),
                                         int(diagonal), C.byref(p), C.byref(o), int(seed),
                                         int(segment), C.byref(rep)))
    return _report(rep)


# ------------------------------------------------------------------ flat files
def write_pdb_flat(xyz, chr_name, path):
    x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    _check(_lib().pmmc_write_pdb_flat(x.ctypes.data_as(_f32p), len(x), chr_name.encode(),
                                      str(path).encode()), "pmmc_write_pdb_flat")


def write_cif_flat(xyz, genomic_pos, chr_name, path):
    x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    g = _i32(genomic_pos)
    _check(_lib().pmmc_write_cif_flat(x.ctypes.data_as(_f32p), g.ctypes.data_as(_i32p), len(g),
                                      len(x), chr_name.encode(), str(path).encode()),
           "pmmc_write_cif_flat")


def heat_write(h, path, header=True, zero_diagonal=False, as_integers=False):
    a = np.ascontiguousarray(h, np.float32)
    n = a.shape[0] if a.size else 0
    _check(_lib().pmmc_heat_write(a.ctypes.data_as(_f32p), n, str(path).encode(), int(header),
                                  int(zero_diagonal), int(as_integers)), "pmmc_heat_write")


def heat_read(path):
    """pmmc_heat_read: (matrix, start or None, resolution or None, diagonal size)."""
    n = C.c_int32()
    st, res, diag = C.c_int32(-1), C.c_int32(-1), C.c_int32()
    _check(_lib().pmmc_heat_read(str(path).encode(), None, 0, C.byref(n), C.byref(st), C.byref(res),
                                 None), "pmmc_heat_read")
    out = np.empty((n.value, n.value), np.float32)
    _check(_lib().pmmc_heat_read(str(path).encode(), out.ctypes.data_as(_f32p), out.size, C.byref(n),
                                 C.byref(st), C.byref(res), C.byref(diag)), "pmmc_heat_read")
    return out, (st.value if st.value != -1 else None), (res.value if res.value != -1 else None), diag.value
