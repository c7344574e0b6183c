// On-disk formats around the cluster tree (include/pmmc_formats.h): HCM model
// files (both layouts), PDB and mmCIF structure files, .heat heatmaps.
//
// Plain host C++ (g++, -ffp-contract=off).  Every number is written with the
// C library's printf conversion the reference uses for it ("%f" of a float,
// "%8.3f", "%.3f", ...) and read with its scanf conversion, and the float
// steps of the structure writers (centring, scaling) round where the
// reference's do, so files are byte-identical to the reference's for the same
// model.  The record layouts are the reference's, because the files are the
// interface (other tools read them).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "levels_host.hpp"
#include "pmmc_formats.h"

using namespace pmmc_host;

namespace {

struct File {
  FILE *f = nullptr;
  File(const char *path, const char *mode) : f(path ? std::fopen(path, mode) : nullptr) {}
  ~File() {
    if (f) std::fclose(f);
  }
  // close and report a write failure
  bool close() {
    FILE *g = f;
    f = nullptr;
    return g && std::ferror(g) == 0 && std::fclose(g) == 0;
  }
};

pmmc_status open_fail(const char *path) {
  return fail(std::string("cannot open ") + (path ? path : "(null)"), PMMC_ERR_INVALID);
}

pmmc_status write_fail(const char *path) {
  return fail(std::string("write failed: ") + (path ? path : "(null)"), PMMC_ERR_INVALID);
}

// ---------------------------------------------------------------- HCM records

// Cluster::toFile (Cluster.hpp:175-181)
void cluster_record(FILE *f, const TCluster &c) {
  std::fprintf(f, "%d %d %c %f %f %f %zu", c.start, c.end, c.orientation, c.pos[0], c.pos[1],
               c.pos[2], c.children.size());
  for (int32_t ch : c.children) std::fprintf(f, " %d", ch);
  std::fprintf(f, "\n");
}

// Cluster::toFilePreviousFormat (Cluster.hpp:183-189): genomic position first
void cluster_record_previous(FILE *f, const TCluster &c) {
  std::fprintf(f, "%d %d %d %f %f %f %zu", (c.start + c.end) / 2, c.start, c.end, c.pos[0],
               c.pos[1], c.pos[2], c.children.size());
  for (int32_t ch : c.children) std::fprintf(f, " %d", ch);
  std::fprintf(f, "\n");
}

// InteractionArc::toFile (InteractionArc.hpp:122-125)
void arc_record(FILE *f, const TChr &c, size_t i) {
  const auto &a = c.arcs[i];
  const auto &x = c.arc_info[i];
  std::fprintf(f, "%d %d %d %d %d %d %d\n", a[0], a[1], x[0], x[1], c.scores[i], x[2], x[3]);
}

// HierarchicalChromosome::toFile(FILE *) (HierarchicalChromosome.cu:530-552)
void write_hcm_new(FILE *f, const pmmc_tree &t) {
  std::fprintf(f, "%d %d %d %d\n", (int)t.cl.size(), (int)t.factors.size(), -1, (int)t.chrs.size());
  for (const std::string &s : t.factors) std::fprintf(f, "%s ", s.c_str());
  std::fprintf(f, "\n");
  for (const TCluster &c : t.cl) cluster_record(f, c);
  for (const TChr &c : t.chrs) {
    std::fprintf(f, "%s %d %d\n", c.name.c_str(), c.root, (int)c.arcs.size());
    for (size_t i = 0; i < c.arcs.size(); ++i) arc_record(f, c, i);
  }
}

// HierarchicalChromosome::toFilePreviousFormat (HierarchicalChromosome.cu:554-593):
// the first chromosome's root and arcs only
void write_hcm_previous(FILE *f, const pmmc_tree &t) {
  const TChr &c = t.chrs[0];
  std::fprintf(f, "%zu %zu %d %zu\n", t.cl.size(), c.arcs.size(), c.root, t.factors.size());
  for (const std::string &s : t.factors) std::fprintf(f, "%s ", s.c_str());
  std::fprintf(f, "\n");
  for (const TCluster &k : t.cl) cluster_record_previous(f, k);
  for (size_t i = 0; i < c.arcs.size(); ++i) {
    const auto &x = c.arc_info[i];
    std::fprintf(f, "%d %d %d %d %d %d 0 0\n", c.arcs[i][0], c.arcs[i][1], c.scores[i], x[3], x[0],
                 x[1]);
  }
}

bool scan(FILE *f, const char *fmt, int *v) { return std::fscanf(f, fmt, v) == 1; }

// vector_insert_unique (common.cpp:442-445)
void add_unique(std::vector<int32_t> &v, int32_t x) {
  if (std::find(v.begin(), v.end(), x) == v.end()) v.push_back(x);
}

// arcs of a chromosome onto its clusters' arc lists, as the readers do
bool attach_arcs(pmmc_tree &t, const TChr &c) {
  for (size_t j = 0; j < c.arcs.size(); ++j) {
    const int32_t a = c.arcs[j][0], b = c.arcs[j][1];
    if (a == -1 || b == -1) continue;
    if (a < 0 || b < 0 || a >= (int32_t)t.cl.size() || b >= (int32_t)t.cl.size()) return false;
    add_unique(t.cl[a].arcs, (int32_t)j);
    add_unique(t.cl[b].arcs, (int32_t)j);
  }
  return true;
}

// HierarchicalChromosome::fromFile(FILE *) (HierarchicalChromosome.cu:619-674)
// with Cluster::fromFile (Cluster.hpp:191-210) and InteractionArc::fromFile
bool read_hcm_new(FILE *f, pmmc_tree &t) {
  int n, n_factors, tmp, n_chrs;
  if (std::fscanf(f, "%d %d %d %d", &n, &n_factors, &tmp, &n_chrs) != 4 || n < 0 || n_factors < 0 ||
      n_chrs < 0)
    return false;
  char buf[4096];
  for (int i = 0; i < n_factors; ++i) {
    if (std::fscanf(f, "%4095s", buf) != 1) return false;
    t.factors.push_back(buf);
  }
  // records are appended as they parse (a header count is not trusted with
  // an allocation of its size)
  t.cl.reserve(std::min(n, 1 << 20));
  for (int i = 0; i < n; ++i) {
    TCluster &c = t.cl.emplace_back();
    int st, en, cnt;
    char o;
    float x, y, z;
    if (std::fscanf(f, "%d %d %c %f %f %f %d", &st, &en, &o, &x, &y, &z, &cnt) != 7 || cnt < 0)
      return false;
    for (int k = 0; k < cnt; ++k) {
      int ch;
      if (!scan(f, "%d", &ch)) return false;
      c.children.push_back(ch);
    }
    c.start = st;
    c.end = en;
    c.genomic_pos = (st + en) / 2;
    c.orientation = o;
    c.pos[0] = x;
    c.pos[1] = y;
    c.pos[2] = z;
  }
  for (int i = 0; i < n_chrs; ++i) {
    TChr c;
    int root, n_arcs;
    if (std::fscanf(f, "%4095s %d %d", buf, &root, &n_arcs) != 3 || n_arcs < 0) return false;
    c.name = buf;
    c.root = root;
    c.first = 0;
    for (int j = 0; j < n_arcs; ++j) {
      int a[7];
      if (std::fscanf(f, "%d %d %d %d %d %d %d", &a[0], &a[1], &a[2], &a[3], &a[4], &a[5], &a[6]) != 7)
        return false;
      c.arcs.push_back({a[0], a[1]});
      c.scores.push_back(a[4]);
      c.arc_info.push_back({a[2], a[3], a[5], a[6]});
    }
    if (!attach_arcs(t, c)) return false;
    t.chrs.push_back(std::move(c));
  }
  return true;
}

// HierarchicalChromosome::fromFilePreviousFormat (HierarchicalChromosome.cu:676-740):
// one chromosome named "chr"; arcs as "start end score factor genomic_start
// genomic_end _ _" with eff_score = score
bool read_hcm_previous(FILE *f, pmmc_tree &t) {
  int n, n_arcs, root, n_factors;
  if (std::fscanf(f, "%d %d %d %d", &n, &n_arcs, &root, &n_factors) != 4 || n < 0 || n_arcs < 0 ||
      n_factors < 0)
    return false;
  char buf[4096];
  for (int i = 0; i < n_factors; ++i) {
    if (std::fscanf(f, "%4095s", buf) != 1) return false;
    t.factors.push_back(buf);
  }
  t.cl.reserve(std::min(n, 1 << 20));
  for (int i = 0; i < n; ++i) {
    TCluster &c = t.cl.emplace_back();
    int gpos, st, en, cnt;
    float x, y, z;
    if (std::fscanf(f, "%d %d %d %f %f %f %d", &gpos, &st, &en, &x, &y, &z, &cnt) != 7 || cnt < 0)
      return false;
    for (int k = 0; k < cnt; ++k) {
      int ch;
      if (!scan(f, "%d", &ch)) return false;
      c.children.push_back(ch);
    }
    c.start = st;
    c.end = en;
    c.genomic_pos = gpos;
    c.pos[0] = x;
    c.pos[1] = y;
    c.pos[2] = z;
  }
  TChr c;
  c.name = "chr";
  c.root = root;
  c.first = 0;
  for (int j = 0; j < n_arcs; ++j) {
    int a[8];
    if (std::fscanf(f, "%d %d %d %d", &a[0], &a[1], &a[2], &a[3]) != 4 ||
        std::fscanf(f, "%d %d %d %d", &a[4], &a[5], &a[6], &a[7]) != 4)
      return false;
    c.arcs.push_back({a[0], a[1]});
    c.scores.push_back(a[2]);
    c.arc_info.push_back({a[4], a[5], a[2], a[3]});
  }
  if (!attach_arcs(t, c)) return false;
  t.chrs.push_back(std::move(c));
  return true;
}

// ---------------------------------------------------------------- levels

// HierarchicalChromosome::useTopLevel + levelDown (HierarchicalChromosome.cu:
// 211-257): `downs` steps from the root, each replacing clusters by their
// children (childless ones stay), then sorting by genomic position
std::vector<int32_t> model_level(const pmmc_tree &t, const TChr &c, int downs) {
  std::vector<int32_t> cur;
  if (c.root >= 0) cur.push_back(c.root);
  for (int d = 0; d < downs; ++d) {
    std::vector<int32_t> tmp;
    for (int32_t k : cur) {
      const TCluster &cl = t.cl[k];
      if (cl.children.empty())
        tmp.push_back(k);
      else
        tmp.insert(tmp.end(), cl.children.begin(), cl.children.end());
    }
    std::sort(tmp.begin(), tmp.end(),
              [&](int a, int b) { return t.cl[a].genomic_pos < t.cl[b].genomic_pos; });
    cur.swap(tmp);
  }
  return cur;
}

bool valid_model(const pmmc_tree &t) {
  const int32_t n = (int32_t)t.cl.size();
  for (const TCluster &c : t.cl)
    for (int32_t ch : c.children)
      if (ch < 0 || ch >= n) return false;
  for (const TChr &c : t.chrs)
    if (c.root < -1 || c.root >= n) return false;
  return true;
}

// ---------------------------------------------------------------- PDB

constexpr int kPdbMaxAtoms = 99999;
constexpr int kPdbMaxResid = 9999;

// chain letter of chromosome i (PdbWriter.hpp:80-88)
char pdb_chain(int i) {
  if (i < 26) return static_cast<char>('A' + i);
  if (i < 52) return static_cast<char>('a' + (i - 26));
  if (i < 62) return static_cast<char>('0' + (i - 52));
  return 'Z';
}

// residue numbers past 9,999 wrap (PdbWriter.hpp:183-185)
int wrap_resid(int r) { return r > kPdbMaxResid ? r % kPdbMaxResid + 1 : r; }

struct Bead {
  int32_t cluster, chr, factor;
};

// beads of every chromosome at the level, in order, with the factor of the
// first of the cluster's arcs that has one (PdbWriter.hpp:136-166)
std::vector<Bead> collect_beads(const pmmc_tree &t, int level, bool with_factor) {
  std::vector<Bead> beads;
  for (size_t ci = 0; ci < t.chrs.size(); ++ci) {
    const TChr &c = t.chrs[ci];
    for (int32_t idx : model_level(t, c, level < 0 ? 5 : level)) {
      if (idx < 0 || idx >= (int32_t)t.cl.size()) continue;
      int32_t factor = -1;
      if (with_factor)
        for (int32_t ai : t.cl[idx].arcs)
          if (ai >= 0 && ai < (int32_t)c.arc_info.size() && c.arc_info[ai][3] >= 0) {
            factor = c.arc_info[ai][3];
            break;
          }
      beads.push_back({idx, (int32_t)ci, factor});
    }
  }
  return beads;
}

// centre and scale to +-50 (PdbWriter.hpp:218-236, CifWriter.hpp:226-241):
// float sums in bead order, the mean by a float division
template <class Div>
void normalisation(const pmmc_tree &t, const std::vector<Bead> &beads, Div divide, float &cx,
                   float &cy, float &cz, float &scale) {
  cx = cy = cz = 0.0f;
  for (const Bead &b : beads) {
    const float *p = t.cl[b.cluster].pos;
    cx += p[0];
    cy += p[1];
    cz += p[2];
  }
  divide(cx, cy, cz);
  float max_dist = 1.0f;
  for (const Bead &b : beads) {
    const float *p = t.cl[b.cluster].pos;
    const float dx = p[0] - cx, dy = p[1] - cy, dz = p[2] - cz;
    const float d = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (d > max_dist) max_dist = d;
  }
  scale = 50.0f / max_dist;
}

// PdbWriter::write (PdbWriter.hpp:90-294)
void write_pdb(FILE *f, const pmmc_tree &t, int level) {
  std::fprintf(f, "HEADER    CHROMOSOME STRUCTURE                        cudaMMC\n");
  std::fprintf(f, "TITLE     3D GENOME STRUCTURE FROM CUDA-MMC\n");
  std::fprintf(f, "REMARK 999 GENOMIC_PDB CONVERTED_FROM_HCM\n");
  for (size_t i = 0; i < t.factors.size(); ++i)
    std::fprintf(f, "REMARK 999 FACTOR %d %s\n", (int)i, t.factors[i].c_str());
  for (const TChr &c : t.chrs) std::fprintf(f, "REMARK 999 CHROMOSOME %s\n", c.name.c_str());
  std::fprintf(f, "REMARK 999 BFACTOR_ANNOTATION NONE=0.00 CTCF=10.00 ENHANCER=20.00 "
                  "PROMOTER=30.00 EPIGENETIC=40.00\n");
  const std::vector<Bead> beads = collect_beads(t, level, true);
  // one RESIDUE record per bead: genomic position, level, factor
  for (size_t i = 0; i < beads.size(); ++i) {
    const TCluster &c = t.cl[beads[i].cluster];
    std::fprintf(f, "REMARK 999 RESIDUE %d GENOMICPOS %d LEVEL %d FACTOR %d\n",
                 wrap_resid((int)(i + 1)), c.genomic_pos, c.level, beads[i].factor);
  }
  // one HELIX record per non-empty chain
  {
    int serial = 1, at = 0;
    for (size_t ci = 0; ci < t.chrs.size(); ++ci) {
      const int len = (int)model_level(t, t.chrs[ci], level < 0 ? 5 : level).size();
      if (len < 1) continue;
      const char id = pdb_chain((int)ci);
      std::fprintf(f, "HELIX  %3d %3d GLY %c %4d  GLY %c %4d  1\n", serial, serial, id,
                   wrap_resid(at + 1), id, wrap_resid(at + len));
      at += len;
      ++serial;
    }
  }
  float cx, cy, cz, scale;
  const int count = (int)beads.size();
  normalisation(t, beads,
                [&](float &x, float &y, float &z) {
                  if (count > 0) {
                    x /= count;
                    y /= count;
                    z /= count;
                  }
                },
                cx, cy, cz, scale);
  int serial = 1, prev = -1;
  for (size_t i = 0; i < beads.size(); ++i) {
    const Bead &b = beads[i];
    if (b.chr != prev && prev >= 0) {
      std::fprintf(f, "TER   %5d      GLY %c%4d\n", std::min(serial, kPdbMaxAtoms), pdb_chain(prev),
                   wrap_resid((int)i));
      ++serial;
    }
    prev = b.chr;
    const TCluster &c = t.cl[b.cluster];
    const float nx = (c.pos[0] - cx) * scale, ny = (c.pos[1] - cy) * scale,
                nz = (c.pos[2] - cz) * scale;
    const float bfactor = (float)c.annotation * 10.0f;
    std::fprintf(f, "ATOM  %5d  CA  GLY %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f          C \n",
                 std::min(serial, kPdbMaxAtoms), pdb_chain(b.chr), wrap_resid((int)(i + 1)), nx, ny,
                 nz, (float)c.level, bfactor);
    ++serial;
  }
  if (!beads.empty())
    std::fprintf(f, "TER   %5d      GLY %c%4d\n", std::min(serial, kPdbMaxAtoms),
                 pdb_chain(beads.back().chr), wrap_resid((int)beads.size()));
  std::fprintf(f, "END\n");
}

// ---------------------------------------------------------------- mmCIF

// CifWriter.hpp:66-77
void cif_header(FILE *f, const char *name) {
  std::fprintf(f, "data_%s\n", name);
  std::fprintf(f, "#\n");
  std::fprintf(f, "_entry.id %s\n", name);
  std::fprintf(f, "#\n");
  std::fprintf(f, "_audit_conform.dict_name       mmcif_pdbx.dic\n");
  std::fprintf(f, "_audit_conform.dict_version    5.296\n");
  std::fprintf(f, "_audit_conform.dict_location   "
                  "http://mmcif.pdb.org/dictionaries/ascii/mmcif_pdbx.dic\n");
  std::fprintf(f, "#\n");
}

// CifWriter.hpp:79-92
void cif_genomic(FILE *f, const std::vector<std::string> &factors,
                 const std::vector<std::string> &chrs) {
  std::fprintf(f, "# GENOMIC_CIF\n");
  for (size_t i = 0; i < factors.size(); ++i)
    std::fprintf(f, "# FACTOR %d %s\n", (int)i, factors[i].c_str());
  for (const std::string &c : chrs) std::fprintf(f, "# CHROMOSOME %s\n", c.c_str());
  std::fprintf(f, "#\n");
}

// CifWriter.hpp:94-113
void cif_atom_site(FILE *f) {
  static const char *const cols[] = {
      "group_PDB",    "id",           "type_symbol",      "label_atom_id",   "label_alt_id",
      "label_comp_id", "label_asym_id", "label_entity_id", "label_seq_id",   "pdbx_PDB_ins_code",
      "Cartn_x",      "Cartn_y",      "Cartn_z",          "occupancy",       "B_iso_or_equiv",
      "pdbx_auth_seq_id", "auth_asym_id"};
  std::fprintf(f, "loop_\n");
  for (const char *c : cols) std::fprintf(f, "_atom_site.%s\n", c);
}

// chain id of chromosome i: A-Z, then two letters (CifWriter.hpp:115-124)
std::string cif_chain(int i) {
  if (i < 26) return std::string(1, static_cast<char>('A' + i));
  return std::string(1, static_cast<char>('A' + (i / 26 - 1))) +
         std::string(1, static_cast<char>('A' + (i % 26)));
}

// a coordinate through the PDB column ("%8.3f" into 8 characters) and back
// (CifWriter.hpp:255-261)
float via_pdb_column(float v) {
  char b[9] = {};
  std::snprintf(b, 9, "%8.3f", (double)v);
  float r = 0.0f;
  std::sscanf(b, "%f", &r);
  return r;
}

// CifWriter::write (CifWriter.hpp:126-298)
void write_cif(FILE *f, const pmmc_tree &t, int level, const char *name) {
  cif_header(f, name);
  std::vector<std::string> chrs;
  for (const TChr &c : t.chrs) chrs.push_back(c.name);
  cif_genomic(f, t.factors, chrs);
  const std::vector<Bead> beads = collect_beads(t, level, false);
  struct Chain {
    std::string id;
    int len;
  };
  std::vector<Chain> chains;
  {
    size_t at = 0;
    for (size_t ci = 0; ci < t.chrs.size(); ++ci) {
      size_t e = at;
      while (e < beads.size() && beads[e].chr == (int32_t)ci) ++e;
      if (e > at) chains.push_back({cif_chain((int)ci), (int)(e - at)});
      at = e;
    }
  }
  static const char *const conf[] = {
      "conf_type_id", "id", "pdbx_PDB_helix_id", "beg_label_comp_id", "beg_label_asym_id",
      "beg_label_seq_id", "pdbx_beg_PDB_ins_code", "end_label_comp_id", "end_label_asym_id",
      "end_label_seq_id", "pdbx_end_PDB_ins_code", "beg_auth_comp_id", "beg_auth_asym_id",
      "beg_auth_seq_id", "end_auth_comp_id", "end_auth_asym_id", "end_auth_seq_id",
      "pdbx_PDB_helix_class", "details", "pdbx_PDB_helix_length"};
  std::fprintf(f, "loop_\n");
  for (const char *c : conf) std::fprintf(f, "_struct_conf.%s\n", c);
  for (size_t h = 0; h < chains.size(); ++h) {
    const char *id = chains[h].id.c_str();
    const int len = chains[h].len;
    std::fprintf(f, "HELX_P HELX_P%d H%d GLY %s %d ? GLY %s %d ? GLY %s %d GLY %s %d 1 ? %d\n",
                 (int)(h + 1), (int)(h + 1), id, 1, id, len, id, 1, id, len, len);
  }
  std::fprintf(f, "#\n");
  std::fprintf(f, "loop_\n");
  std::fprintf(f, "_entity_poly_seq.entity_id\n");
  std::fprintf(f, "_entity_poly_seq.num\n");
  std::fprintf(f, "_entity_poly_seq.mon_id\n");
  std::fprintf(f, "_entity_poly_seq.hetero\n");
  for (size_t h = 0; h < chains.size(); ++h)
    for (int s = 1; s <= chains[h].len; ++s) std::fprintf(f, "%d %d GLY n\n", (int)(h + 1), s);
  std::fprintf(f, "#\n");
  float cx, cy, cz, scale;
  const size_t count = beads.size();
  normalisation(t, beads,
                [&](float &x, float &y, float &z) {
                  if (count > 0) {
                    x /= count;
                    y /= count;
                    z /= count;
                  }
                },
                cx, cy, cz, scale);
  cif_atom_site(f);
  int atom = 1, entity = 1, prev = -1, seq = 1;
  for (const Bead &b : beads) {
    if (b.chr != prev) {
      seq = 1;
      if (prev >= 0) ++entity;
      prev = b.chr;
    }
    const TCluster &c = t.cl[b.cluster];
    const float rx = via_pdb_column((c.pos[0] - cx) * scale);
    const float ry = via_pdb_column((c.pos[1] - cy) * scale);
    const float rz = via_pdb_column((c.pos[2] - cz) * scale);
    const std::string id = cif_chain(b.chr);
    const float mb = (float)c.genomic_pos / 1e6f;
    static const char *const atoms[3][2] = {{"N", "N"}, {"C", "CA"}, {"C", "C"}};
    for (const auto &a : atoms) {
      std::fprintf(f, "ATOM %d %s %s . GLY %s %d %d ? %.3f %.3f %.3f 1.00 %.2f %d %s\n", atom,
                   a[0], a[1], id.c_str(), entity, seq, rx, ry, rz, mb, c.genomic_pos,
                   t.chrs[b.chr].name.c_str());
      ++atom;
    }
    ++seq;
  }
  std::fprintf(f, "#\n");
}

}  // namespace

// ---------------------------------------------------------------- model fields

extern "C" pmmc_status pmmc_tree_set_factors(pmmc_tree *t, int32_t n, const char *const *names) {
  if (!t || n < 0 || (n > 0 && !names)) return fail("pmmc_tree_set_factors: bad arguments", PMMC_ERR_INVALID);
  std::vector<std::string> v;
  for (int32_t i = 0; i < n; ++i) {
    if (!names[i]) return fail("pmmc_tree_set_factors: null name", PMMC_ERR_INVALID);
    v.push_back(names[i]);
  }
  t->factors.swap(v);
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_tree_set_arc_info(pmmc_tree *t, int32_t chr, int32_t n,
                                              const pmmc_arc *arcs) {
  if (!t || chr < 0 || chr >= (int32_t)t->chrs.size() || n < 0 || (n > 0 && !arcs))
    return fail("pmmc_tree_set_arc_info: bad arguments", PMMC_ERR_INVALID);
  TChr &c = t->chrs[chr];
  if ((size_t)n != c.arcs.size())
    return fail("pmmc_tree_set_arc_info: arc count differs from the tree's", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < n; ++i)
    if (arcs[i].score != c.scores[i])
      return fail("pmmc_tree_set_arc_info: arc score differs from the tree's", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < n; ++i)
    c.arc_info[i] = {arcs[i].genomic_start, arcs[i].genomic_end, arcs[i].eff_score, arcs[i].factor};
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_tree_set_annotation(pmmc_tree *t, int32_t n, const int32_t *ids,
                                                const int32_t *types) {
  if (!t || n < 0 || (n > 0 && (!ids || !types)))
    return fail("pmmc_tree_set_annotation: bad arguments", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < n; ++i)
    if (!valid_cluster(t, ids[i])) return fail("pmmc_tree_set_annotation: bad cluster", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < n; ++i) t->cl[ids[i]].annotation = types[i];
  return PMMC_OK;
}

// ---------------------------------------------------------------- HCM

extern "C" pmmc_status pmmc_tree_write_hcm(const pmmc_tree *t, const char *path, int32_t format) {
  if (!t || !path || (format != PMMC_HCM_NEW && format != PMMC_HCM_PREVIOUS))
    return fail("pmmc_tree_write_hcm: bad arguments", PMMC_ERR_INVALID);
  if (format == PMMC_HCM_PREVIOUS && t->chrs.empty())
    return fail("pmmc_tree_write_hcm: the previous format needs a chromosome", PMMC_ERR_INVALID);
  File f(path, "w");
  if (!f.f) return open_fail(path);
  if (format == PMMC_HCM_NEW)
    write_hcm_new(f.f, *t);
  else
    write_hcm_previous(f.f, *t);
  return f.close() ? PMMC_OK : write_fail(path);
}

extern "C" pmmc_status pmmc_tree_read_hcm(const char *path, int32_t format, pmmc_tree **out) {
  if (!path || !out || format < PMMC_HCM_PREVIOUS || format > PMMC_HCM_DETECT)
    return fail("pmmc_tree_read_hcm: bad arguments", PMMC_ERR_INVALID);
  File f(path, "r");
  if (!f.f) return open_fail(path);
  if (format == PMMC_HCM_DETECT) {
    // HierarchicalChromosome::fromFile(string) (HierarchicalChromosome.cu:595-617)
    int n = 0, second = 0, tmp = 0, fourth = 0;
    if (std::fscanf(f.f, "%d %d %d %d", &n, &second, &tmp, &fourth) != 4)
      return fail(std::string("pmmc_tree_read_hcm: bad header in ") + path, PMMC_ERR_INVALID);
    std::fseek(f.f, 0, SEEK_SET);
    format = (n > 100 && second < 5) ? PMMC_HCM_NEW : PMMC_HCM_PREVIOUS;
  }
  pmmc_tree *t = new pmmc_tree();
  bool ok = false;
  try {  // record counts come from the file: no exception leaves the C ABI
    ok = format == PMMC_HCM_NEW ? read_hcm_new(f.f, *t) : read_hcm_previous(f.f, *t);
  } catch (const std::exception &) {
    delete t;
    return fail(std::string("pmmc_tree_read_hcm: host memory exhausted reading ") + path, PMMC_ERR_NOMEM);
  }
  if (!ok || !valid_model(*t)) {
    delete t;
    return fail(std::string("pmmc_tree_read_hcm: malformed file ") + path, PMMC_ERR_INVALID);
  }
  *out = t;
  return PMMC_OK;
}

// ---------------------------------------------------------------- PDB / mmCIF

extern "C" pmmc_status pmmc_tree_write_pdb(const pmmc_tree *t, int32_t level, const char *path) {
  if (!t || !path) return fail("pmmc_tree_write_pdb: bad arguments", PMMC_ERR_INVALID);
  if (!valid_model(*t)) return fail("pmmc_tree_write_pdb: child or root index out of range", PMMC_ERR_INVALID);
  File f(path, "w");
  if (!f.f) return open_fail(path);
  write_pdb(f.f, *t, level);
  return f.close() ? PMMC_OK : write_fail(path);
}

extern "C" pmmc_status pmmc_tree_write_cif(const pmmc_tree *t, int32_t level, const char *data_name,
                                           const char *path) {
  if (!t || !path) return fail("pmmc_tree_write_cif: bad arguments", PMMC_ERR_INVALID);
  if (!valid_model(*t)) return fail("pmmc_tree_write_cif: child or root index out of range", PMMC_ERR_INVALID);
  File f(path, "w");
  if (!f.f) return open_fail(path);
  write_cif(f.f, *t, level, data_name ? data_name : "cudaMMC");
  return f.close() ? PMMC_OK : write_fail(path);
}

// PdbWriter::writeFlat (PdbWriter.hpp:296-335)
extern "C" pmmc_status pmmc_write_pdb_flat(const float *xyz, int32_t n, const char *chr,
                                           const char *path) {
  if (n < 0 || (n > 0 && !xyz) || !path) return fail("pmmc_write_pdb_flat: bad arguments", PMMC_ERR_INVALID);
  (void)chr;  // the reference writes chain A whatever the name
  File f(path, "w");
  if (!f.f) return open_fail(path);
  std::fprintf(f.f, "HEADER    CHROMOSOME STRUCTURE                        cudaMMC\n");
  std::fprintf(f.f, "TITLE     3D GENOME STRUCTURE FROM CUDA-MMC\n");
  for (int32_t i = 0; i < n; ++i)
    std::fprintf(f.f, "ATOM  %5d  CA  ALA A%4d    %8.3f%8.3f%8.3f%6.2f%6.2f          C \n",
                 std::min(i + 1, kPdbMaxAtoms), wrap_resid(i + 1), xyz[3 * i], xyz[3 * i + 1],
                 xyz[3 * i + 2], 1.00f, 0.00f);
  std::fprintf(f.f, "TER   %5d      ALA A%4d\n", std::min(n + 1, kPdbMaxAtoms), std::min(n, kPdbMaxResid));
  std::fprintf(f.f, "END\n");
  return f.close() ? PMMC_OK : write_fail(path);
}

// CifWriter::writeFlat (CifWriter.hpp:300-333)
extern "C" pmmc_status pmmc_write_cif_flat(const float *xyz, const int32_t *genomic_pos, int32_t n_pos,
                                           int32_t n, const char *chr, const char *path) {
  if (n < 0 || (n > 0 && !xyz) || n_pos < 0 || (n_pos > 0 && !genomic_pos) || !path || !chr)
    return fail("pmmc_write_cif_flat: bad arguments", PMMC_ERR_INVALID);
  File f(path, "w");
  if (!f.f) return open_fail(path);
  cif_header(f.f, "cudaMMC");
  cif_genomic(f.f, {}, {chr});
  cif_atom_site(f.f);
  for (int32_t i = 0; i < n; ++i) {
    const int g = i < n_pos ? genomic_pos[i] : i;
    std::fprintf(f.f, "ATOM %d C CA . GLY A 1 %d ? %.3f %.3f %.3f 1.00 %.2f %d %s\n", i + 1, i + 1,
                 xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], (float)g / 1e6f, g, chr);
  }
  std::fprintf(f.f, "#\n");
  return f.close() ? PMMC_OK : write_fail(path);
}

// ---------------------------------------------------------------- .heat

// Heatmap::toFile (Heatmap.hpp:679-700)
extern "C" pmmc_status pmmc_heat_write(const float *h, int32_t n, const char *path, int32_t header,
                                       int32_t zero_diagonal, int32_t as_integers) {
  if (n < 0 || (n > 0 && !h) || !path) return fail("pmmc_heat_write: bad arguments", PMMC_ERR_INVALID);
  File f(path, "w");
  if (!f.f) return open_fail(path);
  if (header) std::fprintf(f.f, "%zu\n", (size_t)n);
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t l = 0; l < n; ++l) {
      float v = h[(size_t)i * n + l];
      if (i == l && zero_diagonal) v = 0.0f;
      if (as_integers)
        std::fprintf(f.f, "%d ", (int)v);
      else
        std::fprintf(f.f, "%f ", v);
    }
    std::fprintf(f.f, "\n");
  }
  return f.close() ? PMMC_OK : write_fail(path);
}

// Heatmap::fromFile(string) (Heatmap.hpp:702-732): the header's field count
// is one plus the number of blanks and tabs in its first 63 characters
// (countWords, common.cpp:426-434)
extern "C" pmmc_status pmmc_heat_read(const char *path, float *out, int64_t cap, int32_t *n,
                                      int32_t *start, int32_t *resolution, int32_t *diag) {
  if (!path || !n) return fail("pmmc_heat_read: bad arguments", PMMC_ERR_INVALID);
  File f(path, "r");
  if (!f.f) return open_fail(path);
  char line[64];
  if (!std::fgets(line, 64, f.f)) return fail(std::string("pmmc_heat_read: empty file ") + path, PMMC_ERR_INVALID);
  int words = 1;
  for (const char *p = line; *p; ++p)
    if (*p == ' ' || *p == '\t') ++words;
  std::fseek(f.f, 0, SEEK_SET);
  int size = 0, st = 0, res = 0;
  if (words == 1) {
    if (std::fscanf(f.f, "%d", &size) != 1) return fail("pmmc_heat_read: bad header", PMMC_ERR_INVALID);
  } else if (words == 3) {
    if (std::fscanf(f.f, "%d %d %d", &size, &st, &res) != 3)
      return fail("pmmc_heat_read: bad header", PMMC_ERR_INVALID);
  } else {
    return fail(std::string("pmmc_heat_read: unrecognized heatmap header in ") + path, PMMC_ERR_INVALID);
  }
  if (size < 0) return fail("pmmc_heat_read: negative size", PMMC_ERR_INVALID);
  // Heatmap::init (Heatmap.hpp:298-301), called after the header is parsed,
  // resets start and resolution: the header's values do not survive
  if (start) *start = -1;
  if (resolution) *resolution = 0;
  *n = size;
  const int64_t need = (int64_t)size * size;
  if (!out) return PMMC_OK;
  if (cap < need) return fail("pmmc_heat_read: output too small", PMMC_ERR_INVALID);
  for (int64_t k = 0; k < need; ++k)
    if (std::fscanf(f.f, "%f", &out[k]) != 1)
      return fail(std::string("pmmc_heat_read: short file ") + path, PMMC_ERR_INVALID);
  if (diag) *diag = diagonal_size(out, size);
  return PMMC_OK;
}
