"""The oracle's annealer (orc_anneal_chain, CPU-preset schedules) agrees with
the reference's own CPU annealers over 16 seeds within the bands stated in
tests/band_common.py.  This pins the restated annealing loop, its schedule
and stop rules against the reference code itself (statistically: the two use
different random streams, Philox here and glibc rand() there)."""
import concurrent.futures as cf
import os

import pytest

import band_common as B
from oracle import binding as O
from oracle import metrics as OM


@pytest.mark.parametrize("level", [1, 2, 3])
def test_oracle_annealing_within_band_of_reference(level):
    cases = {r["seed"]: B.case(level, r["seed"]) for r in B.REF[str(level)]}
    probs = {s: O.Problem(c[0]) for s, c in cases.items()}

    def one(job):
        s, chain = job
        prob, sched, truth = cases[s]
        pos, st = probs[s].anneal(O.schedule_from(sched), 1 + s, s, chain)
        rec = {"energy": st["energy"], "moves": st["moves"]}
        rec.update(B.quality(level, pos, prob, truth, OM))
        return job, rec

    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = dict(ex.map(one, B.jobs(level, B.K_CPU[level])))
    stats = B.compare(level, results, B.K_CPU[level])
    print(level, stats)
