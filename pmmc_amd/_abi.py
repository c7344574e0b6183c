"""ctypes mirror of include/pmmc_hip.h and include/pmmc_levels.h (field order and types match the C
structs exactly) and the loader of libpmmc_hip.so.

The library is the product: if it is missing this module raises, there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmmc_hip.so")

LEVEL_HEATMAP, LEVEL_ARCS, LEVEL_SMOOTH = 1, 2, 3
PRESET_CPU, PRESET_GPU = 0, 1
STOP_FIRST_FAIL, STOP_MISSES_RESET, STOP_MISSES_ACCUM = 0, 1, 2
STATUS = {0: "PMMC_OK", 1: "PMMC_ERR_INVALID", 2: "PMMC_ERR_HIP", 3: "PMMC_ERR_NOMEM",
          4: "PMMC_ERR_NODEVICE", 5: "PMMC_ERR_BUSY"}
STOP_REASONS = {0: "running", 1: "rule", 2: "small", 3: "cap", 4: "guard", 5: "nofree",
                6: "nonfinite"}

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)


class Schedule(C.Structure):
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

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LevelSettings(C.Structure):
    _fields_ = [
        ("max_temp", C.c_double), ("dt_temp", C.c_double), ("improvement", C.c_double),
        ("tjs", C.c_float), ("tjc", C.c_float),
        ("min_successes", C.c_int32), ("steps", C.c_int32), ("milestone_fails", C.c_int32),
        ("sphere_radius", C.c_float), ("use2d", C.c_int32),
    ]


class HeatmapProblem(C.Structure):
    _fields_ = [("n", C.c_int32), ("pos", f32p), ("fixed", u8p), ("dist", f32p),
                ("diagonal", C.c_int32), ("boundaries", i32p), ("n_boundaries", C.c_int32),
                ("skip_below", C.c_float)]


class ArcsProblem(C.Structure):
    _fields_ = [("n", C.c_int32), ("pos", f32p), ("fixed", u8p), ("exp_dist", f32p),
                ("n_loops", C.c_int32), ("loop_pairs", i32p), ("loop_params", f32p),
                ("k_stretch", C.c_float), ("k_squeeze", C.c_float)]


class SmoothProblem(C.Structure):
    _fields_ = [("n", C.c_int32), ("pos", f32p), ("fixed", u8p), ("dist_to_next", f32p),
                ("k_stretch", C.c_float), ("k_squeeze", C.c_float), ("k_angular", C.c_float),
                ("w_dist", C.c_float), ("w_angle", C.c_float)]


class Run(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("segment", C.c_uint32), ("chain0", C.c_uint32),
                ("n_chains", C.c_int32)]


class SingletonInput(C.Structure):
    _fields_ = [("n_chr", C.c_int32), ("cluster_offsets", i32p), ("cluster_start", i32p),
                ("cluster_end", i32p), ("n_records", C.c_int64), ("chr1", i32p), ("start1", i32p),
                ("end1", i32p), ("chr2", i32p), ("start2", i32p), ("end2", i32p), ("score", i32p),
                ("region_start", C.c_int32), ("region_end", C.c_int32), ("centromere", i32p),
                ("n_long", C.c_int32), ("long_chr", i32p), ("long_start", i32p), ("long_end", i32p),
                ("long_score", i32p), ("long_scale", C.c_float), ("long_power", C.c_float),
                ("diag", C.c_int32)]


class SingletonCounts(C.Structure):
    _fields_ = [("added", C.c_int64), ("diagonal", C.c_int64), ("bad_chr", C.c_int64),
                ("kernel_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ProbeResult(C.Structure):
    _fields_ = [("moves_per_s", C.c_double), ("read_gbps", C.c_double), ("write_gbps", C.c_double),
                ("stream_read_gbps", C.c_double), ("lds_gbps", C.c_double), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("energy", C.c_double), ("energy_resummed", C.c_double),
                ("best_energy", C.c_double), ("moves", C.c_int64), ("accepted", C.c_int64),
                ("best_chain_moves", C.c_int64), ("best_chain_accepted", C.c_int64),
                ("best_chain", C.c_int32), ("n_chains", C.c_int32), ("milestones", C.c_int32),
                ("stop_reason", C.c_int32), ("kernel_ms", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["stop_reason"] = STOP_REASONS.get(d["stop_reason"], d["stop_reason"])
        return d


class Job(C.Structure):
    _fields_ = [("level", C.c_int32), ("problem", C.c_void_p), ("schedule", Schedule),
                ("run", Run), ("pos_out", f32p), ("stats", C.POINTER(Stats))]


class BatchOpts(C.Structure):
    _fields_ = [("blocks_per_step", C.c_int32), ("steps_per_graph", C.c_int32),
                ("use_graph", C.c_int32), ("chains_per_wave", C.c_int32), ("tuned", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SweepSchedule(C.Structure):
    _fields_ = [("T0", C.c_double), ("dt", C.c_double), ("improvement", C.c_double),
                ("small_energy", C.c_double),
                ("tjs", C.c_float), ("tjc", C.c_float), ("step", C.c_float),
                ("step_decay", C.c_float), ("sphere_radius", C.c_float),
                ("inner", C.c_int32), ("proposers", C.c_int32), ("fails_threshold", C.c_int32),
                ("max_iters", C.c_int32), ("use2d", C.c_int32)]


class Arc(C.Structure):
    _fields_ = [(k, C.c_int32) for k in
                ("start", "end", "genomic_start", "genomic_end", "score", "eff_score", "factor")]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_jobs", C.c_int32),
                ("blocks_per_step", C.c_int32), ("launch_ms", C.c_float), ("streams", C.c_int32)]


# name -> (restype, argtypes); every symbol declared in include/pmmc_hip.h
class LevelParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "count_a", "count_scale", "count_shift", "count_base", "heat_scale", "heat_power",
        "heat_scale_inter", "heat_power_inter", "heat_stretching", "glen_base", "glen_scale",
        "glen_power")] + [("loop_density", C.c_int32)] + [(n, C.c_float) for n in (
        "noise_chr", "noise_segment", "noise_anchor", "noise_subanchor", "noise_anchor_small")] + [
        (n, C.c_int32) for n in ("restarts_chr", "restarts_segment", "restarts_anchor",
                                 "restarts_subanchor", "use_anchor_heatmap")] + [
        ("anchor_heatmap_influence", C.c_float), ("use2d", C.c_int32)]


class Chromosome(C.Structure):
    _fields_ = [("name", C.c_char_p), ("n_anchors", C.c_int32), ("anchor_start", i32p),
                ("anchor_end", i32p), ("orientation", C.c_char_p), ("n_arcs", C.c_int32),
                ("arc_start", i32p), ("arc_end", i32p), ("arc_score", i32p),
                ("n_long_arcs", C.c_int32), ("long_start", i32p), ("long_end", i32p),
                ("long_score", i32p)]


class SegmentRegion(C.Structure):
    _fields_ = [("chr", C.c_char_p), ("start", C.c_int32), ("end", C.c_int32)]


class Cluster(C.Structure):
    _fields_ = [("start", C.c_int32), ("end", C.c_int32), ("genomic_pos", C.c_int32),
                ("base_start", C.c_int32), ("base_end", C.c_int32), ("level", C.c_int32),
                ("parent", C.c_int32), ("n_children", C.c_int32), ("is_fixed", C.c_int32),
                ("orientation", C.c_int32), ("pos", C.c_float * 3), ("dist_to_next", C.c_double)]


class LevelReport(C.Structure):
    _fields_ = [("n_blocks", C.c_int32), ("n_jobs", C.c_int32), ("moves", C.c_int64),
                ("accepted", C.c_int64), ("score_anchor", C.c_double), ("score_smooth", C.c_double),
                ("ms_anchor", C.c_double), ("ms_smooth", C.c_double),
                ("score_global", C.c_double), ("ms_global", C.c_double),
                ("stress_iterations", C.c_int32), ("reinterpolated", C.c_int32)]


class StressProblem(C.Structure):
    _fields_ = [("n", C.c_int32), ("pos", f32p), ("backbone", f32p), ("n_pairs", C.c_int32),
                ("pair_i", i32p), ("pair_j", i32p), ("pair_target", f32p),
                ("k_stretch", C.c_float), ("k_squeeze", C.c_float), ("k_angular", C.c_float)]


class StressOpts(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("lr", C.c_float), ("in_order_sums", C.c_int32)]


class StressStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("stop", C.c_int32), ("lr", C.c_float),
                ("exact_sums", C.c_int32), ("total", C.c_double), ("backbone", C.c_double),
                ("pairs", C.c_double), ("angular", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class ArcsLevelOpts(C.Structure):
    _fields_ = [("arcs", LevelSettings), ("smooth", LevelSettings), ("arcs_preset", C.c_int32),
                ("smooth_preset", C.c_int32), ("k_stretch_arcs", C.c_float),
                ("k_squeeze_arcs", C.c_float), ("k_stretch", C.c_float), ("k_squeeze", C.c_float),
                ("k_angular", C.c_float), ("w_dist", C.c_float), ("w_angle", C.c_float),
                ("chains_per_restart", C.c_int32), ("densify", C.c_int32),
                ("global_refine", C.c_int32), ("stress", StressOpts),
                ("one_call_per_job", C.c_int32)]


class HeatmapLevelOpts(C.Structure):
    _fields_ = [("heatmap", LevelSettings), ("preset", C.c_int32), ("chains_per_restart", C.c_int32),
                ("one_call_per_job", C.c_int32)]


SIGNATURES = {
    "pmmc_schedule_preset": (C.c_int, [C.c_int, C.c_int, C.POINTER(LevelSettings), C.c_float,
                                       C.POINTER(Schedule)]),
    "pmmc_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "pmmc_ctx_destroy": (None, [C.c_void_p]),
    "pmmc_ctx_reserve": (C.c_int, [C.c_void_p, C.c_size_t]),
    "pmmc_last_error": (C.c_char_p, [C.c_void_p]),
    "pmmc_ctx_cu_count": (C.c_int, [C.c_void_p]),
    "pmmc_set_gpu_seed": (None, [C.c_void_p, C.c_uint64]),
    "pmmc_pick_shift": (C.c_int32, [C.c_int32, C.c_int32]),
    "pmmc_default_chains": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "pmmc_heatmap_anneal": (C.c_int, [C.c_void_p, C.POINTER(HeatmapProblem), C.POINTER(Schedule),
                                      C.POINTER(Run), f32p, C.POINTER(Stats)]),
    "pmmc_arcs_anneal": (C.c_int, [C.c_void_p, C.POINTER(ArcsProblem), C.POINTER(Schedule),
                                   C.POINTER(Run), f32p, C.POINTER(Stats)]),
    "pmmc_smooth_anneal": (C.c_int, [C.c_void_p, C.POINTER(SmoothProblem), C.POINTER(Schedule),
                                     C.POINTER(Run), f32p, C.POINTER(Stats)]),
    "pmmc_anneal_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(Schedule),
                                      C.POINTER(Run), f32p, C.POINTER(Stats)]),
    "pmmc_heatmap_sweep": (C.c_int, [C.c_void_p, C.POINTER(HeatmapProblem),
                                     C.POINTER(SweepSchedule), C.c_uint64, C.c_uint32, f32p,
                                     C.POINTER(Stats)]),
    "pmmc_delta_energy": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_float, C.c_int32, i32p,
                                    f32p, f32p]),
    "pmmc_mark_arcs": (C.c_int, [C.c_void_p, i32p, i32p, C.c_int32, i32p, i32p, C.c_int32, i32p,
                                 i32p]),
    "pmmc_group_arcs": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p, i32p, i32p, i32p,
                                  C.POINTER(Arc), i32p, i32p]),
    "pmmc_anneal_batch": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_int32, C.POINTER(BatchOpts)]),
    "pmmc_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]),
    "pmmc_server_destroy": (None, [C.c_void_p]),
    "pmmc_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Job), C.POINTER(C.c_uint64)]),
    "pmmc_server_submit_many": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_int32,
                                          C.POINTER(C.c_uint64)]),
    "pmmc_server_step": (C.c_int, [C.c_void_p, i32p, i32p, i32p]),
    "pmmc_server_done": (C.c_int, [C.c_void_p, C.c_uint64]),
    "pmmc_singleton_heatmap": (C.c_int, [C.c_void_p, C.POINTER(SingletonInput), f32p,
                                         C.POINTER(SingletonCounts)]),
    "pmmc_probe_move_pattern": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32, C.c_int32, C.POINTER(ProbeResult)]),
    "pmmc_stress_opts_default": (None, [C.POINTER(StressOpts)]),
    "pmmc_stress_refine": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(StressProblem),
                                     C.POINTER(StressOpts), C.POINTER(f32p),
                                     C.POINTER(StressStats)]),
    "pmmc_probe_acosf": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32,
                                   C.POINTER(C.c_uint64)]),
    "pmmc_spline_count": (C.c_int64, [C.c_int64, C.c_int64, C.c_int32]),
    "pmmc_spline_positions": (C.c_int, [C.c_void_p, C.c_int32, i32p, f32p, C.c_int32, C.c_int32,
                                        C.c_int64, C.c_int32, C.c_int64, f32p]),
    # include/pmmc_levels.h
    "pmmc_level_params_default": (None, [C.POINTER(LevelParams)]),
    "pmmc_host_last_error": (C.c_char_p, []),
    "pmmc_freq_to_distance": (C.c_float, [C.c_int32, C.POINTER(LevelParams)]),
    "pmmc_genomic_length_to_distance": (C.c_double, [C.c_int32, C.POINTER(LevelParams)]),
    "pmmc_normalize_heatmap": (C.c_int, [f32p, C.c_int32, f32p]),
    "pmmc_normalize_heatmap_diagonal_total": (C.c_int, [f32p, C.c_int32, C.c_float]),
    "pmmc_normalize_heatmap_inter": (C.c_int, [f32p, C.c_int32, C.c_int32, i32p, C.c_float]),
    "pmmc_distance_heatmap": (C.c_int, [f32p, C.c_int32, C.c_int32, C.c_int32,
                                        C.POINTER(LevelParams), f32p, f32p]),
    "pmmc_tree_create": (C.c_int, [C.c_int32, C.POINTER(Chromosome), C.c_int32,
                                   C.POINTER(SegmentRegion), C.POINTER(C.c_void_p)]),
    "pmmc_tree_destroy": (None, [C.c_void_p]),
    "pmmc_tree_size": (C.c_int32, [C.c_void_p]),
    "pmmc_tree_n_chromosomes": (C.c_int32, [C.c_void_p]),
    "pmmc_tree_get": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(Cluster)]),
    "pmmc_tree_children": (C.c_int, [C.c_void_p, C.c_int32, i32p, C.c_int32, i32p]),
    "pmmc_tree_chromosome": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p]),
    "pmmc_tree_arcs": (C.c_int, [C.c_void_p, C.c_int32, i32p, C.c_int32, i32p]),
    "pmmc_tree_level": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, i32p, C.c_int32, i32p]),
    "pmmc_tree_set_positions": (C.c_int, [C.c_void_p, C.c_int32, i32p, f32p]),
    "pmmc_tree_get_positions": (C.c_int, [C.c_void_p, C.c_int32, i32p, f32p]),
    "pmmc_tree_set_fixed": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p]),
    "pmmc_tree_densify": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, i32p]),
    "pmmc_tree_set_dist_to_next": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(LevelParams),
                                             f32p, f32p]),
    "pmmc_tree_expected_distances": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32,
                                               C.POINTER(LevelParams), f32p, f32p]),
    "pmmc_tree_stress_refine": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32,
                                          C.POINTER(LevelParams), C.c_float, C.c_float, C.c_float,
                                          C.POINTER(StressOpts), C.POINTER(StressStats)]),
    "pmmc_arcs_level_opts_default": (None, [C.POINTER(ArcsLevelOpts)]),
    "pmmc_heatmap_level_opts_default": (None, [C.POINTER(HeatmapLevelOpts)]),
    "pmmc_arcs_level": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(LevelParams),
                                  C.POINTER(ArcsLevelOpts), C.c_uint64, C.c_uint32,
                                  C.POINTER(LevelReport)]),
    "pmmc_heatmap_level": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, f32p, C.c_int32,
                                     C.POINTER(LevelParams), C.POINTER(HeatmapLevelOpts), C.c_uint64,
                                     C.c_uint32, C.POINTER(LevelReport)]),
    # include/pmmc_formats.h
    "pmmc_tree_set_factors": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_char_p)]),
    "pmmc_tree_set_arc_info": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(Arc)]),
    "pmmc_tree_set_annotation": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p]),
    "pmmc_tree_write_hcm": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32]),
    "pmmc_tree_read_hcm": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "pmmc_tree_write_pdb": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p]),
    "pmmc_tree_write_cif": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, C.c_char_p]),
    "pmmc_write_pdb_flat": (C.c_int, [f32p, C.c_int32, C.c_char_p, C.c_char_p]),
    "pmmc_write_cif_flat": (C.c_int, [f32p, i32p, C.c_int32, C.c_int32, C.c_char_p, C.c_char_p]),
    "pmmc_heat_write": (C.c_int, [f32p, C.c_int32, C.c_char_p, C.c_int32, C.c_int32, C.c_int32]),
    "pmmc_heat_read": (C.c_int, [C.c_char_p, f32p, C.c_int64, i32p, i32p, i32p, i32p]),
    "pmmc_pearson": (C.c_float, [f32p, f32p, C.c_int64]),
    "pmmc_distance_correlation": (C.c_float, [f32p, f32p, C.c_int32, C.c_int32]),
    "pmmc_distance_map": (C.c_int, [f32p, C.c_int32, f32p]),
    "pmmc_scc": (C.c_float, [f32p, f32p, C.c_int32, C.c_int32]),
    "pmmc_rmsd": (C.c_float, [f32p, f32p, C.c_int32]),
}

_lib = None


def load() -> C.CDLL:
    """Load libpmmc_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} not found: build it with `python -m pmmc_amd.build` "
            "(the HIP library is required; there is no fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
