"""Writes tests/golden/formats_ref.json: what the reference's own record
writers and .heat reader/writer produce, for the on-disk format tests.

The outputs come from oracle/_ref/levels_ref (oracle/ref_levels_driver.cpp),
compiled from the reference sources in place by oracle/Makefile:
  records   Cluster::toFile, Cluster::toFilePreviousFormat (Cluster.hpp:175-189)
            and InteractionArc::toFile (InteractionArc.hpp:122-125) -- the HCM
            records -- on the clusters and arcs of a densified two-chromosome
            tree with awkward coordinates (negative zero, 1e-7, 1e10, large
            and negative values) and arcs with and without factors
  heat      Heatmap::toFile (Heatmap.hpp:679-700) for every flag combination,
            and Heatmap::fromFile (:702-732) of the 1-field and 3-field
            headers
  flat      PdbWriter::writeFlat (PdbWriter.hpp:296-335) and
            CifWriter::writeFlat (CifWriter.hpp:300-333, with the header,
            genomic-metadata and atom-site helpers :66-124 it shares with
            the model writer) on special values (signed zero, rounding
            halves, NaN and infinities, values wider than the PDB columns),
            genomic positions shorter than the bead list, and a
            100,050-bead chain past the PDB resSeq wrap (9,999) and serial
            cap (99,999); that case's inputs come from a seed and its
            outputs are kept as SHA-256 digests with their first and last
            lines
The HCM framing lines (HierarchicalChromosome::toFile) and the model forms
of the PDB / mmCIF writers (PdbWriter::write, CifWriter::write) need a
HierarchicalChromosome, whose members live in a CUDA source file the image
cannot compile; their tests check structure and round trips instead.  Run
where the reference sources exist:
    make -C oracle ref && python tests/golden/make_formats_fixtures.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
BIN = os.path.join(ROOT, "oracle", "_ref", "levels_ref")
HEADER = "0 " * 12


def hx(a):
    return [float(v).hex() for v in np.asarray(a, np.float32).reshape(-1)]


def records_case():
    """Clusters and arcs of a product tree (two chromosomes, densified)."""
    from pmmc_amd import tree as T
    rng = np.random.default_rng(5)
    chrs = []
    for c, n in enumerate((9, 7)):
        starts = 10_000 + np.cumsum(rng.integers(3000, 9000, size=n))
        anchors = [(int(s), int(s + rng.integers(200, 900))) for s in starts]
        arcs = [(i, min(n - 1, i + int(rng.integers(1, 3))), int(rng.integers(2, 40)))
                for i in range(0, n - 1, 2)]
        chrs.append(dict(name=f"chr{c + 1}", anchors=anchors, arcs=arcs,
                         orient=["L", "R", "N"] * 3))
    splits = []
    for c in chrs:
        lo, hi = c["anchors"][0][0] - 100, c["anchors"][-1][1] + 100
        splits += [(c["name"], lo, (lo + hi) // 2), (c["name"], (lo + hi) // 2, hi)]
    t = T.Tree(chrs, splits)
    for ci in range(2):
        for ib in t.level(ci, 2):
            t.densify(ib, 2)
    n = len(t)
    pos = rng.normal(0, 30.0, size=(n, 3)).astype(np.float32)
    special = np.array([-0.0, 1e-7, 1e10, -123456.789, 0.5, -1e-30, 3.4e38, 7.0], np.float32)
    pos.reshape(-1)[:special.size] = special
    t.set_positions(list(range(n)), pos)
    cl = t.clusters()
    clusters = []
    for i, k in enumerate(cl):
        clusters.append(dict(start=k.start, end=k.end, orientation=chr(k.orientation),
                             pos=hx(pos[i]), children=t.children(i)))
    arcs = []
    for ci in range(2):
        for (a, b), (_, _, score) in zip(t.arcs(ci), chrs[ci]["arcs"]):
            arcs.append([a, b, int(rng.integers(1000, 9_000_000)), int(rng.integers(1000, 9_000_000)),
                         score, int(rng.integers(1, 60)), int(rng.integers(-1, 3))])
    return dict(chrs=chrs, splits=splits, pos=hx(pos), clusters=clusters, arcs=arcs)


def run_records(case):
    text = "records " + HEADER + "\n" + f"{len(case['clusters'])}\n"
    for c in case["clusters"]:
        text += (f"{c['start']} {c['end']} {c['orientation']} " + " ".join(c["pos"]) +
                 f" {len(c['children'])} " + " ".join(map(str, c["children"])) + "\n")
    text += f"{len(case['arcs'])}\n" + "".join(" ".join(map(str, a)) + "\n" for a in case["arcs"])
    r = subprocess.run([BIN], input=text, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out, cur = {}, None
    for ln in r.stdout.splitlines(keepends=True):
        if ln.startswith("@@"):
            cur = ln.strip()[2:]
            out[cur] = ""
        elif cur is not None:
            out[cur] += ln
    return {k: v for k, v in out.items() if k != "end"}


def run_heat(mat, flags, read_text):
    with tempfile.TemporaryDirectory() as d:
        w, rd = os.path.join(d, "w.heat"), os.path.join(d, "r.heat")
        with open(rd, "w") as f:
            f.write(read_text)
        n = mat.shape[0]
        text = f"heat {HEADER}\n{w} {n} {flags[0]} {flags[1]} {flags[2]}\n"
        text += " ".join(hx(mat)) + f"\n{rd}\n"
        r = subprocess.run([BIN], input=text, capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        written = open(w).read()
    lines = r.stdout.splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.startswith("@@heat"))
    size, start, res, diag = (int(x) for x in lines[k].split()[1:])
    vals = [float.fromhex(x) for ln in lines[k + 1:k + 1 + size] for x in ln.split()]
    return written, dict(size=size, start=start, resolution=res, diagonal=diag, values=hx(vals))


FLAT_SPECIAL = [[-0.0, 0.0005, -0.0005], [0.00049999997, 1234.5675, -1234.5675],
                [1e10, -99999.9, 1e-30], [float("nan"), float("inf"), float("-inf")],
                [3.4e38, -3.4e38, 7.0], [12.25, -0.125, 99.9995], [1.0, 2.0, 3.0]]


def flat_inputs(case):
    """Positions [n, 3] float32 and genomic positions of a flat case."""
    if case["kind"] == "special":
        pos = np.array(FLAT_SPECIAL, np.float32)
        return pos, np.array([0, 5, -5, 2_000_000_000, 123_456_789], np.int32)
    rng = np.random.default_rng(case["seed"])
    n = case["n"]
    pos = rng.normal(0.0, case["scale"], size=(n, 3)).astype(np.float32)
    gpos = np.cumsum(rng.integers(1, 50_000, size=case["n_pos"])).astype(np.int32)
    return pos, gpos


def run_flat(pos, gpos, chr_name):
    with tempfile.TemporaryDirectory() as d:
        pdb, cif = os.path.join(d, "f.pdb"), os.path.join(d, "f.cif")
        text = f"flat {HEADER}\n{pdb} {cif} {chr_name} {len(pos)} {len(gpos)}\n"
        text += " ".join(hx(pos)) + "\n" + " ".join(str(int(g)) for g in gpos) + "\n"
        r = subprocess.run([BIN], input=text, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr
        with open(pdb, "rb") as f:
            a = f.read()
        with open(cif, "rb") as f:
            b = f.read()
    return a, b


def summary(data):
    lines = data.decode().splitlines(keepends=True)
    return dict(sha256=hashlib.sha256(data).hexdigest(), size=len(data), head=lines[:12],
                tail=lines[-4:])


def flat_cases():
    cases = [dict(kind="special", chr="chr1"),
             dict(kind="random", seed=11, n=200, n_pos=200, scale=30.0, chr="chrX"),
             dict(kind="random", seed=12, n=40, n_pos=25, scale=800.0, chr="chr22"),
             dict(kind="random", seed=13, n=100_050, n_pos=100_050, scale=20.0, chr="chr2")]
    out = []
    for c in cases:
        pos, gpos = flat_inputs(c)
        a, b = run_flat(pos, gpos, c["chr"])
        if len(pos) <= 1000:
            c = dict(c, pdb=a.decode(), cif=b.decode())
        else:
            c = dict(c, pdb_summary=summary(a), cif_summary=summary(b))
        out.append(c)
    return out


def main():
    fx = dict(records=records_case())
    fx["flat"] = flat_cases()
    fx["records"]["expected"] = run_records(fx["records"])
    rng = np.random.default_rng(9)
    heat = []
    mats = [rng.uniform(0, 50, size=(6, 6)).astype(np.float32),
            np.array([[0, 0, 3.5, 1e-6], [0, 0, 0, -2.25], [3.5, 0, 0, 0], [1e-6, -2.25, 0, 7e5]],
                     np.float32),
            np.zeros((1, 1), np.float32)]
    for m in mats:
        m = (m + m.T) / 2 if m.shape[0] > 1 else m
        for flags in [(1, 0, 0), (0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)]:
            # read back a 3-field header file with the same values
            rtext = f"{m.shape[0]} 150000 25000\n" + "".join(
                " ".join(f"{float(v):.6f}" for v in row) + " \n" for row in m)
            written, read = run_heat(m, flags, rtext)
            heat.append(dict(matrix=hx(m), n=m.shape[0], flags=list(flags), written=written,
                             read_text=rtext, read=read))
    # the reference reading its own 1-field-header output
    m = mats[0]
    m = (m + m.T) / 2
    w1, _ = run_heat(m, (1, 0, 0), "1\n0\n")
    _, r1 = run_heat(m, (1, 0, 0), w1)
    heat.append(dict(matrix=hx(m), n=m.shape[0], flags=[1, 0, 0], written=w1, read_text=w1, read=r1))
    fx["heat"] = heat
    with open(os.path.join(HERE, "formats_ref.json"), "w") as f:
        json.dump(fx, f, separators=(",", ":"))
    print("records:", {k: len(v.splitlines()) for k, v in fx["records"]["expected"].items()},
          "heat cases:", len(heat))


if __name__ == "__main__":
    main()
