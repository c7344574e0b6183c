"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/*.h, the ctypes mirrors have the C layouts, and
the schedule presets reproduce the reference annealers' constants."""
import ctypes as C
import glob
import os
import re
import subprocess
import tempfile

import pytest

from pmmc_amd import _abi as A
from pmmc_amd.api import schedule_preset, settings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_functions():
    txt = "".join(open(h).read() for h in HEADERS)
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pmmc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(A.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(A.SIGNATURES), set(names) ^ set(A.SIGNATURES)


STRUCTS = {
    "pmmc_schedule": A.Schedule, "pmmc_level_settings": A.LevelSettings,
    "pmmc_heatmap_problem": A.HeatmapProblem, "pmmc_arcs_problem": A.ArcsProblem,
    "pmmc_smooth_problem": A.SmoothProblem, "pmmc_run": A.Run, "pmmc_anneal_stats": A.Stats,
    "pmmc_job": A.Job, "pmmc_batch_opts": A.BatchOpts, "pmmc_sweep_schedule": A.SweepSchedule,
    "pmmc_arc": A.Arc, "pmmc_server_opts": A.ServerOpts,
    "pmmc_level_params": A.LevelParams, "pmmc_chromosome": A.Chromosome,
    "pmmc_segment_region": A.SegmentRegion, "pmmc_cluster": A.Cluster,
    "pmmc_level_report": A.LevelReport, "pmmc_arcs_level_opts": A.ArcsLevelOpts,
    "pmmc_heatmap_level_opts": A.HeatmapLevelOpts, "pmmc_probe_result": A.ProbeResult,
    "pmmc_singleton_input": A.SingletonInput, "pmmc_singleton_counts": A.SingletonCounts,
}


def test_ctypes_layouts_match_header():
    prog = ['#include <stdio.h>', '#include <stddef.h>'] + [f'#include "{h}"' for h in HEADERS]
    prog.append("int main(void){")
    for cname, py in STRUCTS.items():
        prog.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            prog.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    prog.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "t.c"), os.path.join(d, "t")
        open(src, "w").write("\n".join(prog))
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{os.path.join(ROOT, 'include')}",
                        src, "-o", exe], check=True)
        out = subprocess.run([exe], stdout=subprocess.PIPE, text=True, check=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    for cname, py in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"


def test_no_device_gives_status_not_crash():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = A.load()
    h = C.c_void_p()
    assert lib.pmmc_ctx_create(0, C.byref(h)) == 4     # PMMC_ERR_NODEVICE


def test_presets_follow_reference_constants():
    s = settings(max_temp=5.0, dt_temp=0.9999, improvement=0.99, tjs=50, tjc=20,
                 min_successes=10, steps=50000, milestone_fails=3)
    hc = schedule_preset(1, 0, s, 2.0)     # LooperSolver::MonteCarloHeatmap
    assert hc.step == pytest.approx(1.0) and hc.block == 50000 and hc.stop_rule == 0
    assert hc.small_energy == 1e-6 and hc.cool_per_block == 0 and hc.guard_energy == 1e-6
    hg = schedule_preset(1, 1, s, 2.0)     # MonteCarloHeatmapKernel
    assert hg.step == pytest.approx(1.5) and hg.block == 512 and hg.cool_per_block == 1
    assert hg.step_decay == pytest.approx(0.95) and hg.stop_rule == 2 and hg.small_energy == 1e-4
    ac = schedule_preset(2, 0, s, 0.1)     # LooperSolver::MonteCarloArcs
    assert ac.ratio_stop == 0.9999 and ac.small_energy == 1e-5 and ac.positive_guard == 0
    ag = schedule_preset(2, 1, s, 0.1)     # MonteCarloArcsKernel
    assert ag.block == 256 and ag.stop_rule == 1 and ag.positive_guard == 1
    assert ag.max_moves == 50000
    sc = schedule_preset(3, 0, s, 0.1)     # LooperSolver::MonteCarloArcsSmooth
    assert sc.strict_accept == 1 and sc.block == 50000
    sg = schedule_preset(3, 1, s, 0.1)     # MonteCarloSmoothKernelMultiWarp
    assert sg.block == 256 and sg.milestone_period == 50000 and sg.max_moves == 50000
    assert sg.stop_rule == 2
    with pytest.raises(Exception):
        schedule_preset(7, 0, s, 1.0)


def test_pick_shift_rule_matches_oracle():
    from oracle import binding as O
    from pmmc_amd import api
    for level in (1, 2, 3):
        for n in (1, 2, 31, 32, 33, 48, 64, 65, 400):
            P = O.Problem.__new__(O.Problem)
            P.level, P.n = level, n
            assert api.pick_shift(level, n) == P.pick_shift(), (level, n)
    assert api.pick_shift(1, 64) == 0 and api.pick_shift(1, 65) == 6 and api.pick_shift(2, 48) == 6
