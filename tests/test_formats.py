"""On-disk formats (include/pmmc_formats.h): HCM model files, PDB and mmCIF
structure files, .heat heatmaps.  Host code only; no GPU.

* HCM records: every cluster record (new and previous layout) and arc record
  the product writes equals the reference's own Cluster::toFile /
  toFilePreviousFormat / InteractionArc::toFile output for the same model,
  byte for byte (tests/golden/formats_ref.json, from the reference compiled
  in place); the framing lines follow HierarchicalChromosome::toFile
  (HierarchicalChromosome.cu:530-593) -- a CUDA source the image cannot
  compile, so they are checked for structure; both layouts read back to the
  same model, and the reference's layout-detection rule is applied.
* .heat: the written text equals Heatmap::toFile's for every flag
  combination, and reading equals Heatmap::fromFile (values, size, diagonal
  band, start / resolution as its init leaves them), byte / bit exact.
* PDB and mmCIF flat writers: byte-exact against the reference's own
  PdbWriter::writeFlat / CifWriter::writeFlat compiled in place (the mmCIF
  header, genomic-metadata and atom-site helpers included, which the model
  writer shares), on special values, short genomic-position lists and a
  100,050-bead chain (SHA-256 of the reference's output).
* PDB and mmCIF model writers: the reference's PdbWriter::write /
  CifWriter::write need a HierarchicalChromosome, whose members live in a
  CUDA source file the image cannot compile, so their byte parity is
  unpinned; checked instead: the PDB fixed columns parse back to the scaled
  coordinates (centroid 0, farthest bead at 50), levels, annotation
  B-factors, chains, TER / HELIX / RESIDUE records, residue numbers wrapping
  past 9,999, and the mmCIF coordinates equal the PDB ones (the reference
  rounds them through the PDB column).
"""
import hashlib
import importlib.util
import json
import os

import numpy as np
import pytest

from pmmc_amd import tree as T

HERE = os.path.dirname(os.path.abspath(__file__))
FX = json.load(open(os.path.join(HERE, "golden", "formats_ref.json")))


def unhex(v):
    return np.array([float.fromhex(x) for x in v], np.float32)


def records_tree():
    r = FX["records"]
    t = T.Tree(r["chrs"], r["splits"])
    for ci in range(2):
        for ib in t.level(ci, 2):
            t.densify(ib, 2)
    n = len(t)
    t.set_positions(list(range(n)), unhex(r["pos"]).reshape(n, 3))
    k = 0
    for ci in range(2):
        m = len(t.arcs(ci))
        t.set_arc_info(ci, r["arcs"][k:k + m])
        k += m
    t.set_factors(["CTCF", "RNAPII", "H3K27ac"])
    return t


def test_hcm_new_records_match_reference(tmp_path):
    t = records_tree()
    p = tmp_path / "m.hcm"
    t.write_hcm(p, fmt=1)
    lines = open(p).read().splitlines(keepends=True)
    n = len(t)
    exp = FX["records"]["expected"]
    assert lines[0] == f"{n} 3 -1 2\n"
    assert lines[1] == "CTCF RNAPII H3K27ac \n"
    assert "".join(lines[2:2 + n]) == exp["cluster"]
    rest = lines[2 + n:]
    arcs = []
    for ci in range(2):
        first, root = t.chromosome(ci)
        m = len(t.arcs(ci))
        assert rest[0] == f"chr{ci + 1} {root} {m}\n"
        arcs += rest[1:1 + m]
        rest = rest[1 + m:]
    assert rest == []
    assert "".join(arcs) == exp["arc"]


def test_hcm_previous_records_match_reference(tmp_path):
    t = records_tree()
    p = tmp_path / "m.hcm"
    t.write_hcm(p, fmt=0)
    lines = open(p).read().splitlines(keepends=True)
    n = len(t)
    exp = FX["records"]["expected"]
    _, root = t.chromosome(0)
    m = len(t.arcs(0))
    assert lines[0] == f"{n} {m} {root} 3\n"
    assert "".join(lines[2:2 + n]) == exp["cluster_previous"]
    # the first chromosome's arcs: start end score factor genomic_start genomic_end 0 0
    for ln, a in zip(lines[2 + n:], FX["records"]["arcs"][:m]):
        s, e, gs, ge, sc, eff, fac = a
        assert ln == f"{s} {e} {sc} {fac} {gs} {ge} 0 0\n"
    assert len(lines) == 2 + n + m


def test_hcm_round_trip_both_layouts(tmp_path):
    t = records_tree()
    n = len(t)
    for fmt in (1, 0):
        p = tmp_path / f"m{fmt}.hcm"
        t.write_hcm(p, fmt=fmt)
        r = T.Tree.from_hcm(p, fmt=fmt)
        assert len(r) == n
        a, b = t.clusters(), r.clusters()
        for i in range(n):
            assert (a[i].start, a[i].end) == (b[i].start, b[i].end)
            assert a[i].genomic_pos == b[i].genomic_pos
            assert t.children(i) == r.children(i)
            if fmt == 1:
                assert a[i].orientation == b[i].orientation
        # "%f" keeps 6 decimals: the positions come back within half a unit of
        # the 6th decimal plus the float rounding of the parsed value
        got, want = r.positions(list(range(n))), t.positions(list(range(n)))
        small = np.abs(want) < 1e6
        np.testing.assert_allclose(got[small], want[small], atol=5.01e-7, rtol=1.2e-7)
        assert r.n_chromosomes == (2 if fmt == 1 else 1)
        assert r.chromosome(0)[1] == t.chromosome(0)[1]
        assert r.arcs(0) == t.arcs(0)
        # writing the read model again gives the same file
        q = tmp_path / f"again{fmt}.hcm"
        r.write_hcm(q, fmt=fmt)
        if fmt == 1:
            assert open(q).read() == open(p).read()


def test_hcm_detection_rule(tmp_path):
    """More than 100 clusters and fewer than 5 factors: the new layout; a
    small model written in the new layout is taken for the previous one (the
    reference's rule), which fails to parse here instead of reading garbage."""
    t = records_tree()
    p = tmp_path / "small.hcm"
    t.write_hcm(p, fmt=1)
    assert len(t) <= 100
    with pytest.raises(RuntimeError):
        T.Tree.from_hcm(p, fmt=2)
    big = T.Tree([dict(name="chrA", anchors=[(10_000 + 5000 * i, 10_400 + 5000 * i) for i in range(60)],
                       arcs=[(i, i + 1, 5) for i in range(0, 58, 3)])],
                 [("chrA", 9000, 320_000)])
    for ib in big.level(0, 2):
        big.densify(ib, 2)
    assert len(big) > 100
    q = tmp_path / "big.hcm"
    big.write_hcm(q, fmt=1)
    r = T.Tree.from_hcm(q, fmt=2)
    assert len(r) == len(big) and r.n_chromosomes == 1
    q0 = tmp_path / "big0.hcm"
    big.write_hcm(q0, fmt=0)
    r0 = T.Tree.from_hcm(q0, fmt=2)       # previous layout: second field = arcs (>= 5 here)
    assert len(r0) == len(big)


@pytest.mark.parametrize("fmt,header", [(1, "2000000000 1 0 1\nchr1\n"), (0, "2000000000 0 0 1\nchr1\n")])
def test_hcm_header_count_beyond_the_file(tmp_path, fmt, header):
    """A header promising two billion clusters with one record behind it is a
    malformed file: rejected as such, without an allocation of the promised
    size."""
    import resource
    p = tmp_path / "bad.hcm"
    rec = "0 10 N 1 2 3 0\n" if fmt == 1 else "5 0 10 1 2 3 0\n"
    p.write_text(header + rec)
    before = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss
    with pytest.raises(RuntimeError, match="malformed"):
        T.Tree.from_hcm(p, fmt)
    assert resource.getrusage(resource.RUSAGE_SELF).ru_maxrss - before < 512 * 1024  # KiB


@pytest.mark.parametrize("k", range(len(FX["heat"])))
def test_heat_write_and_read_match_reference(tmp_path, k):
    c = FX["heat"][k]
    m = unhex(c["matrix"]).reshape(c["n"], c["n"])
    p = tmp_path / "w.heat"
    hdr, zero, ints = c["flags"]
    T.heat_write(m, p, header=bool(hdr), zero_diagonal=bool(zero), as_integers=bool(ints))
    assert open(p).read() == c["written"]
    q = tmp_path / "r.heat"
    q.write_text(c["read_text"])
    got, start, res, diag = T.heat_read(q)
    r = c["read"]
    assert got.shape == (r["size"], r["size"])
    assert np.array_equal(got.reshape(-1).view(np.uint32), unhex(r["values"]).view(np.uint32))
    assert (start if start is not None else -1, res if res is not None else -1) == (r["start"], r["resolution"])
    assert diag == r["diagonal"]


def test_heat_unrecognised_header(tmp_path):
    q = tmp_path / "bad.heat"
    q.write_text("3 1\n1 2 3\n")
    with pytest.raises(RuntimeError):
        T.heat_read(q)


def pdb_atoms(path):
    """ATOM records by PDB column (1-based columns of the format's spec)."""
    out = []
    for ln in open(path):
        if ln.startswith("ATOM  "):
            out.append(dict(serial=int(ln[6:11]), name=ln[12:16], res=ln[17:20], chain=ln[21],
                            seq=int(ln[22:26]), x=float(ln[30:38]), y=float(ln[38:46]),
                            z=float(ln[46:54]), occ=float(ln[54:60]), b=float(ln[60:66]),
                            elem=ln[76:78], width=len(ln.rstrip("\n"))))
    return out


def lowest_level(t):
    """Each chromosome's leaves sorted by genomic position."""
    cl = t.clusters()
    out = []
    for ci in range(t.n_chromosomes):
        _, root = t.chromosome(ci)
        stack, leaves = [root], []
        while stack:
            k = stack.pop()
            ch = t.children(k)
            if ch:
                stack += ch
            else:
                leaves.append(k)
        out.append(sorted(leaves, key=lambda k: cl[k].genomic_pos))
    return out


def test_pdb_structure(tmp_path):
    t = records_tree()
    n = len(t)
    pos = np.random.default_rng(3).normal(0, 40.0, size=(n, 3)).astype(np.float32)
    t.set_positions(list(range(n)), pos)
    ids = list(range(0, n, 5))
    t.set_annotation(ids, [1 + i % 4 for i in range(len(ids))])
    p = tmp_path / "m.pdb"
    t.write_pdb(p)
    txt = open(p).read().splitlines()
    assert txt[0] == "HEADER    CHROMOSOME STRUCTURE                        cudaMMC"
    assert txt[-1] == "END"
    levels = lowest_level(t)
    beads = [k for lv in levels for k in lv]
    atoms = pdb_atoms(p)
    assert len(atoms) == len(beads)
    cl = t.clusters()
    xyz = np.array([[a["x"], a["y"], a["z"]] for a in atoms])
    P = pos[beads].astype(np.float64)
    c = P.mean(0)
    scale = 50.0 / max(1.0, np.linalg.norm(P - c, axis=1).max())
    np.testing.assert_allclose(xyz, (P - c) * scale, atol=2e-3)
    assert abs(np.linalg.norm(xyz, axis=1).max() - 50.0) < 2e-3
    ann = dict(zip(ids, [1 + i % 4 for i in range(len(ids))]))
    for i, (a, k) in enumerate(zip(atoms, beads)):
        # the element sits at column 77, left-justified, as the reference writes it
        assert a["seq"] == i + 1 and a["name"] == " CA " and a["res"] == "GLY" and a["elem"] == "C "
        assert a["width"] == 78          # the record ends at column 78
        assert a["occ"] == cl[k].level
        assert a["b"] == 10.0 * ann.get(k, 0)
    chains = [a["chain"] for a in atoms]
    assert chains == ["A"] * len(levels[0]) + ["B"] * len(levels[1])
    # serials skip one number at the chain change (the TER record)
    assert atoms[len(levels[0])]["serial"] == len(levels[0]) + 2
    ters = [ln for ln in txt if ln.startswith("TER")]
    helix = [ln for ln in txt if ln.startswith("HELIX")]
    assert len(ters) == 2 and len(helix) == 2
    res = [ln.split() for ln in txt if ln.startswith("REMARK 999 RESIDUE")]
    assert [int(r[5]) for r in res] == [cl[k].genomic_pos for k in beads]
    for lv in levels:
        g = [cl[k].genomic_pos for k in lv]
        assert g == sorted(g)
    assert [ln for ln in txt if "FACTOR 0" in ln][0] == "REMARK 999 FACTOR 0 CTCF"
    # anchors carry the factor of their first arc that has one
    fac = [int(r[9]) for r in res]
    assert any(f >= 0 for f in fac) and any(f == -1 for f in fac)


def test_cif_coordinates_equal_pdb(tmp_path):
    t = records_tree()
    n = len(t)
    pos = np.random.default_rng(4).normal(0, 25.0, size=(n, 3)).astype(np.float32)
    t.set_positions(list(range(n)), pos)
    t.write_pdb(tmp_path / "m.pdb")
    t.write_cif(tmp_path / "m.cif")
    atoms = pdb_atoms(tmp_path / "m.pdb")
    rows = [ln.split() for ln in open(tmp_path / "m.cif") if ln.startswith("ATOM ")]
    assert len(rows) == 3 * len(atoms)
    for i, a in enumerate(atoms):
        for j, (sym, name) in enumerate((("N", "N"), ("C", "CA"), ("C", "C"))):
            r = rows[3 * i + j]
            assert (r[2], r[3]) == (sym, name)
            assert (float(r[10]), float(r[11]), float(r[12])) == (a["x"], a["y"], a["z"])
    txt = open(tmp_path / "m.cif").read()
    assert txt.startswith("data_cudaMMC\n#\n_entry.id cudaMMC\n")
    assert "# GENOMIC_CIF\n# FACTOR 0 CTCF\n" in txt
    assert txt.count("HELX_P HELX_P") == 2


def test_pdb_residue_numbers_wrap(tmp_path):
    xyz = np.zeros((10_003, 3), np.float32)
    xyz[:, 0] = np.arange(10_003) * np.float32(0.5)   # within the 8-character column
    p = tmp_path / "flat.pdb"
    T.write_pdb_flat(xyz, "chr1", p)
    atoms = pdb_atoms(p)
    assert [a["seq"] for a in atoms[9997:10003]] == [9998, 9999, 2, 3, 4, 5]
    assert atoms[-1]["x"] == 5001.0
    assert open(p).read().splitlines()[-2] == "TER   10004      ALA A9999"
    q = tmp_path / "flat.cif"
    T.write_cif_flat(xyz[:3], [100, 2_000_000], "chr7", q)
    rows = [ln.split() for ln in open(q) if ln.startswith("ATOM ")]
    assert [r[-2:] for r in rows] == [["100", "chr7"], ["2000000", "chr7"], ["2", "chr7"]]
    assert rows[1][-3] == "2.00"


def _golden_script():
    path = os.path.join(HERE, "golden", "make_formats_fixtures.py")
    spec = importlib.util.spec_from_file_location("make_formats_fixtures", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("k", range(len(FX["flat"])))
def test_flat_writers_match_reference(tmp_path, k):
    """pmmc_write_pdb_flat / pmmc_write_cif_flat against the reference's
    PdbWriter::writeFlat / CifWriter::writeFlat on the same inputs."""
    case = FX["flat"][k]
    pos, gpos = _golden_script().flat_inputs(case)
    p, q = tmp_path / "f.pdb", tmp_path / "f.cif"
    T.write_pdb_flat(pos, case["chr"], p)
    T.write_cif_flat(pos, gpos, case["chr"], q)
    a, b = p.read_bytes(), q.read_bytes()
    if "pdb" in case:
        assert a.decode() == case["pdb"]
        assert b.decode() == case["cif"]
    else:
        for data, want in ((a, case["pdb_summary"]), (b, case["cif_summary"])):
            lines = data.decode().splitlines(keepends=True)
            assert lines[:len(want["head"])] == want["head"]
            assert lines[-len(want["tail"]):] == want["tail"]
            assert len(data) == want["size"]
            assert hashlib.sha256(data).hexdigest() == want["sha256"]
