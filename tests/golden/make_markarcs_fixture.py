"""Writes tests/golden/markarcs_ref.json: seeded random chromosomes (anchors
and raw PET arcs) and the anchor-indexed arcs that the reference's own
InteractionArcs::markArcs (InteractionArcs.hpp:207-336, built from the
reference sources into oracle/_ref by oracle/Makefile) emits for them.
Run in the container that holds the reference: python tests/golden/make_markarcs_fixture.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import binding as O  # noqa: E402


def chromosome(rng, A, R):
    starts = np.sort(rng.integers(1, 200_000, size=A))
    lens = rng.integers(1, 4000, size=A)
    anc = np.stack([starts, starts + lens], 1)
    anc[rng.integers(0, A)] = (0, 0)            # zero-length anchor
    anc[rng.integers(0, A)] = (starts[0], starts[0])  # length-1 anchor
    anc = anc[np.argsort(anc[:, 0], kind="stable")]
    pick = rng.integers(0, A, size=(R, 2))
    gs = anc[pick[:, 0], 0] + rng.integers(0, 60, R)
    ge = anc[pick[:, 1], 0] + rng.integers(0, 60, R)
    raw = np.stack([gs, ge, rng.integers(1, 30, R), rng.integers(0, 3, R)], 1)
    raw = np.concatenate([raw, raw[: R // 4], [[5, 7, 3, 0]]])   # duplicates, unmatched
    return anc.astype(np.int32), raw.astype(np.int32)


def main():
    rng = np.random.default_rng(20261203)
    chrs = [chromosome(rng, A, R) for A, R in [(30, 120), (60, 300), (15, 40)]]
    ref = O.ref_mark_arcs(chrs)
    if ref is None:
        raise SystemExit("oracle/_ref/markarcs_ref not built (needs the reference sources)")
    out = {"source": "reference InteractionArcs::markArcs (CPU, std::map grouping)",
           "chromosomes": [{"anchors": a.tolist(), "raw": r.tolist(), "arcs": o.tolist()}
                           for (a, r), o in zip(chrs, ref)]}
    with open(os.path.join(os.path.dirname(__file__), "markarcs_ref.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("arcs per chromosome:", [len(o) for o in ref])


if __name__ == "__main__":
    main()
