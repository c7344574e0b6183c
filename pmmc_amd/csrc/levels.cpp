// Host-side rows around the annealing kernels: cluster tree and bead
// indexing, densification, expected-distance inputs (include/pmmc_levels.h).
//
// Plain host C++ (built with g++, -ffp-contract=off): every float operation
// rounds to float exactly where the reference's does, and powf / pow / exp are
// the same libm calls, so results are bit-identical to the reference's host
// code on the same inputs.
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "levels_host.hpp"

namespace pmmc_host {

thread_local std::string g_err;

pmmc_status fail(const std::string &msg, pmmc_status st) {
  g_err = msg;
  return st;
}

}  // namespace pmmc_host

using namespace pmmc_host;

extern "C" const char *pmmc_host_last_error(void) { return g_err.c_str(); }

extern "C" void pmmc_level_params_default(pmmc_level_params *p) {
  // Settings::init (Settings.hpp:395-461)
  p->count_a = 0.5f;
  p->count_scale = 20.0f;
  p->count_shift = 1.0f;
  p->count_base = 0.01f;
  p->heat_scale = 100.0f;
  p->heat_power = -0.333f;
  p->heat_scale_inter = 100.0f;
  p->heat_power_inter = -1.0f;
  p->heat_stretching = 2.0f;
  p->glen_base = 0.0f;
  p->glen_scale = 1.0f;
  p->glen_power = 0.5f;
  p->loop_density = 5;
  p->noise_chr = 1.0f;
  p->noise_segment = 0.1f;
  p->noise_anchor = 0.5f;
  p->noise_subanchor = 0.5f;
  p->noise_anchor_small = 0.05f;
  p->restarts_chr = 2;
  p->restarts_segment = 2;
  p->restarts_anchor = 5;
  p->restarts_subanchor = 5;
  p->use_anchor_heatmap = 0;
  p->anchor_heatmap_influence = 0.5f;
  p->use2d = 0;
}

// ---------------------------------------------------------------- distances

// LooperSolver::freqToDistance (LooperSolver.hpp:3949-3967).  `freq + shift`
// is an int + float sum in float, as there.
extern "C" float pmmc_freq_to_distance(int32_t freq, const pmmc_level_params *p) {
  if (p->count_a < 0.0f) {
    int fi = static_cast<int>(static_cast<float>(freq) + p->count_shift);
    if (fi < 1) fi = 1;
    return static_cast<float>(static_cast<double>(p->count_base) +
                              static_cast<double>(p->count_scale) *
                                  std::pow(static_cast<double>(fi), static_cast<double>(p->count_a)));
  }
  const float x = static_cast<float>(freq) + p->count_shift;
  return static_cast<float>(static_cast<double>(p->count_base) +
                            static_cast<double>(p->count_scale) /
                                std::exp(static_cast<double>(p->count_a) * static_cast<double>(x)));
}

// genomicLengthToDistance (LooperSolver.hpp:3900-3904)
extern "C" double pmmc_genomic_length_to_distance(int32_t length, const pmmc_level_params *p) {
  return static_cast<double>(p->glen_base) +
         static_cast<double>(p->glen_scale) *
             std::pow(length / 1000.0, static_cast<double>(p->glen_power));
}

namespace pmmc_host {

// Heatmap::getDiagonalSize (Heatmap.hpp:356-367)
int diagonal_size(const float *h, int n) {
  for (int w = 0; w < n; ++w)
    for (int i = 0; i + w < n; ++i)
      if (static_cast<double>(h[(size_t)i * n + i + w]) > 1e-6) return w;
  return 0;
}

// Heatmap::getAvg (Heatmap.hpp:440-451): sequential float sum
float heat_avg(const float *h, int n) {
  float ret = 0.0f;
  int cnt = 0;
  const int w = diagonal_size(h, n);
  for (int i = 0; i < n; ++i)
    for (int j = i + w; j < n; ++j) {
      ret += h[(size_t)i * n + j];
      cnt++;
    }
  return ret / static_cast<float>(cnt);
}

// Heatmap::getRange (Heatmap.hpp:427-438): max over the band-free triangle
float heat_max(const float *h, int n) {
  float mx = 0.0f;
  const int w = diagonal_size(h, n);
  for (int i = 0; i < n; ++i)
    for (int j = i + w; j < n; ++j)
      if (h[(size_t)i * n + j] > mx) mx = h[(size_t)i * n + j];
  return mx;
}

}  // namespace pmmc_host

// LooperSolver::createDistanceHeatmap (LooperSolver.hpp:3042-3085)
extern "C" pmmc_status pmmc_distance_heatmap(const float *freq, int32_t n, int32_t diag,
                                             int32_t inter, const pmmc_level_params *p,
                                             float *out, float *avg_out) {
  if (!freq || !out || !p || n < 0)
    return fail("pmmc_distance_heatmap: null pointer or negative size", PMMC_ERR_INVALID);
  const float scale = inter ? p->heat_scale_inter : p->heat_scale;
  const float power = inter ? p->heat_power_inter : p->heat_power;
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      const float v = freq[(size_t)i * n + j];
      float d;
      if (static_cast<double>(v) < 1e-6)
        d = 0.0f;
      else
        d = (std::abs(i - j) < diag) ? -1.0f : scale * powf(v, power);
      out[(size_t)i * n + j] = d;
      out[(size_t)j * n + i] = d;
    }
  const float avg = heat_avg(out, n);
  const float max_dist = avg * p->heat_stretching;
  for (size_t k = 0; k < (size_t)n * n; ++k)
    if (out[k] > max_dist) out[k] = max_dist;
  if (avg_out) *avg_out = avg;
  return PMMC_OK;
}

// LooperSolver::normalizeHeatmap (LooperSolver.hpp:2996-3040)
extern "C" pmmc_status pmmc_normalize_heatmap(const float *in, int32_t n, float *out) {
  if (!in || !out || n < 0)
    return fail("pmmc_normalize_heatmap: null pointer or negative size", PMMC_ERR_INVALID);
  std::vector<float> rows(n);
  float expected = 0.0f;
  for (int i = 0; i < n; ++i) {
    float sum = 0.0f;
    for (int j = 0; j < n; ++j) sum += in[(size_t)i * n + j];
    rows[i] = sum;
    expected += sum;
  }
  expected /= n;
  std::vector<float> h((size_t)n * n);
  for (int i = 0; i < n; ++i) {
    const float mn = rows[i] > 1e-12f ? expected / rows[i] : 0.0f;
    for (int j = 0; j < n; ++j) h[(size_t)i * n + j] = in[(size_t)i * n + j] * mn;
  }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const float v = (h[(size_t)i * n + j] + h[(size_t)j * n + i]) / 2.0f;
      h[(size_t)i * n + j] = v;
      h[(size_t)j * n + i] = v;
    }
  std::copy(h.begin(), h.end(), out);
  return PMMC_OK;
}

// LooperSolver::normalizeHeatmapDiagonalTotal (LooperSolver.hpp:3146-3172)
extern "C" pmmc_status pmmc_normalize_heatmap_diagonal_total(float *heat, int32_t n, float val) {
  if (!heat || n < 0)
    return fail("pmmc_normalize_heatmap_diagonal_total: null pointer or negative size",
                PMMC_ERR_INVALID);
  const int w = diagonal_size(heat, n);
  double avg = 0.0;
  for (int i = 0; i + w < n; ++i) avg += heat[(size_t)i * n + i + w];
  avg /= static_cast<double>(n - w);
  if (!(avg > 1e-12)) return PMMC_OK;
  const float mn = static_cast<float>(val / avg);
  for (size_t k = 0; k < (size_t)n * n; ++k) heat[k] *= mn;
  return PMMC_OK;
}

// LooperSolver::normalizeHeatmapInter (LooperSolver.hpp:3106-3144)
extern "C" pmmc_status pmmc_normalize_heatmap_inter(float *heat, int32_t n, int32_t n_chr,
                                                    const int32_t *sizes, float scale) {
  if (!heat || n < 0 || n_chr < 0 || (n_chr > 0 && !sizes))
    return fail("pmmc_normalize_heatmap_inter: null pointer or negative size", PMMC_ERR_INVALID);
  int64_t total = 0;
  for (int c = 0; c < n_chr; ++c) {
    if (sizes[c] < 0) return fail("pmmc_normalize_heatmap_inter: negative size", PMMC_ERR_INVALID);
    total += sizes[c];
  }
  if (total != n) return fail("pmmc_normalize_heatmap_inter: sizes do not sum to n", PMMC_ERR_INVALID);
  if (n_chr <= 1) return PMMC_OK;
  for (size_t k = 0; k < (size_t)n * n; ++k) heat[k] *= scale;
  int o = 0;
  for (int c = 0; c < n_chr; ++c) {
    for (int i = o; i < o + sizes[c]; ++i)
      for (int j = o; j < o + sizes[c]; ++j) heat[(size_t)i * n + j] /= scale;
    o += sizes[c];
  }
  return PMMC_OK;
}

// ---------------------------------------------------------------- tree

namespace {

int32_t cdiv2(int32_t a) { return a / 2; }  // C++ truncation, as the reference

TCluster make_cluster(int32_t start, int32_t end) {
  // Cluster::Cluster(int, int) (Cluster.hpp:115-136)
  TCluster c;
  c.start = start;
  c.end = end;
  c.genomic_pos = cdiv2(start + end);
  return c;
}

bool contains(const std::vector<int32_t> &v, int32_t x) {
  return std::find(v.begin(), v.end(), x) != v.end();
}

void insert_unique(std::vector<int32_t> &v, int32_t x) {
  if (!contains(v, x)) v.push_back(x);
}

// other end of arc `a` of chromosome c, -1 when the cluster is not on it
int32_t other_end(const TChr &c, int32_t a, int32_t cluster) {
  if (c.arcs[a][0] == cluster) return c.arcs[a][1];
  if (c.arcs[a][1] == cluster) return c.arcs[a][0];
  return -1;
}

// LooperSolver::findGaps (LooperSolver.hpp:2119-2180)
std::vector<int32_t> find_gaps(const pmmc_tree &t, const TChr &c) {
  std::vector<int32_t> gaps;
  const int32_t start = c.first;
  gaps.push_back(start);
  int cnt = 0;
  for (int32_t i = start; i < (int32_t)t.cl.size(); ++i) {
    for (int32_t a : t.cl[i].arcs) {
      const int32_t other = other_end(c, a, i);
      if (other == i) continue;
      if (other > i)
        cnt++;
      else
        cnt--;
    }
    if (cnt == 0) gaps.push_back(i);
  }
  insert_unique(gaps, (int32_t)t.cl.size() - 1);
  if (!t.regions.empty()) {
    for (const TRegion &r : t.regions) {
      if (r.chr != c.name) continue;
      for (int32_t i = start; i < (int32_t)t.cl.size(); ++i)
        if (t.cl[i].end >= r.end) {
          insert_unique(gaps, i);
          break;
        }
    }
    std::sort(gaps.begin(), gaps.end());
    gaps.erase(std::unique(gaps.begin(), gaps.end()), gaps.end());
  }
  return gaps;
}

// LooperSolver::findSplit (LooperSolver.hpp:2185-2269)
bool find_split(const pmmc_tree &t, const TChr &c, const std::vector<int32_t> &gaps,
                std::vector<int32_t> &splits) {
  if (t.cl.size() < gaps.size()) {
    char b[160];
    snprintf(b, sizeof b, "find split: size mismatch, gaps=%zu, clusters=%zu", gaps.size(),
             t.cl.size());
    g_err = b;
    return false;
  }
  if (t.regions.empty()) {
    splits = gaps;
    return true;
  }
  splits.assign(1, gaps[0]);
  int first = -1;
  int last = (int)t.regions.size() - 1;
  for (size_t i = 0; i < t.regions.size(); ++i) {
    if (t.regions[i].chr == c.name) {
      if (first == -1) first = (int)i;
    } else if (first != -1) {
      last = (int)i - 1;
      break;
    }
  }
  if (first == -1) {
    g_err = "no segment split found for " + c.name;
    return false;
  }
  for (int s = first; s < last; ++s) {
    const int32_t boundary = t.regions[s].end;
    int best = -1;
    int best_d = INT_MAX;
    for (size_t i = 1; i + 1 < gaps.size(); ++i) {
      const int d = std::abs(t.cl[gaps[i]].end - boundary);
      if (d < best_d) {
        best_d = d;
        best = (int)i;
      }
    }
    if (best >= 0) insert_unique(splits, gaps[best]);
  }
  std::sort(splits.begin(), splits.end());
  insert_unique(splits, gaps.back());
  return true;
}

// LooperSolver::createTreeChromosome (LooperSolver.hpp:2296-2438)
bool build_chromosome(pmmc_tree &t, TChr &c) {
  const int32_t cs = (int32_t)t.cl.size();
  c.first = cs;
  for (size_t i = 0; i < c.anchors.size(); ++i) {
    TCluster k = make_cluster(c.anchors[i][0], c.anchors[i][1]);
    k.orientation = c.orient.empty() ? 'N' : c.orient[i];
    k.base_start = k.base_end = cs + (int32_t)i;
    k.level = 4;
    t.cl.push_back(k);
  }
  for (size_t i = 0; i < c.arcs.size(); ++i) {
    c.arcs[i][0] += cs;
    c.arcs[i][1] += cs;
    t.cl[c.arcs[i][0]].arcs.push_back((int32_t)i);
    t.cl[c.arcs[i][1]].arcs.push_back((int32_t)i);
  }
  TCluster root;
  const std::vector<int32_t> gaps = find_gaps(t, c);
  std::vector<int32_t> splits;
  if (!find_split(t, c, gaps, splits)) return false;
  size_t next_split = 1;
  int32_t start_gap = (int32_t)t.cl.size();
  for (size_t i = 1; i < gaps.size(); ++i) {
    const int32_t prev_gap = i == 1 ? gaps[i - 1] : gaps[i - 1] + 1;
    const int32_t cur_gap = gaps[i];
    TCluster ib = make_cluster(t.cl[prev_gap].start, t.cl[cur_gap].end);
    ib.base_start = prev_gap;
    ib.base_end = cur_gap;
    ib.level = 3;
    for (int32_t k = prev_gap; k <= cur_gap; ++k) {
      ib.children.push_back(k);
      t.cl[k].parent = (int32_t)t.cl.size();
    }
    t.cl.push_back(ib);
    if (next_split >= splits.size()) {
      g_err = "chromosome " + c.name + ": interaction blocks past the last split";
      return false;
    }
    if (gaps[i] == splits[next_split]) {
      const int32_t end_gap = (int32_t)t.cl.size() - 1;
      TCluster seg = make_cluster(t.cl[start_gap].start, t.cl[end_gap].end);
      seg.base_start = t.cl[start_gap].base_start;
      seg.base_end = t.cl[end_gap].base_end;
      seg.level = 2;
      const int32_t seg_ind = (int32_t)t.cl.size();
      for (int32_t k = start_gap; k <= end_gap; ++k) {
        seg.children.push_back(k);
        t.cl[k].parent = seg_ind;
      }
      root.children.push_back(seg_ind);
      t.cl.push_back(seg);
      start_gap = (int32_t)t.cl.size();
      next_split++;
    }
  }
  if (root.children.empty()) {
    g_err = "chromosome " + c.name + ": no segment could be formed (single anchor with "
            "predefined segments)";
    return false;
  }
  root.level = 1;
  root.start = t.cl[root.children.front()].start;
  root.end = t.cl[root.children.back()].end;
  root.genomic_pos = cdiv2(root.start + root.end);
  root.base_start = t.cl[root.children.front()].base_start;
  root.base_end = t.cl[root.children.back()].base_end;
  for (int32_t ch : root.children) t.cl[ch].parent = (int32_t)t.cl.size();
  c.root = (int32_t)t.cl.size();
  t.cl.push_back(root);
  return true;
}

}  // namespace

extern "C" pmmc_status pmmc_tree_create(int32_t n_chr, const pmmc_chromosome *chrs,
                                        int32_t n_regions, const pmmc_segment_region *regions,
                                        pmmc_tree **out) {
  if (!out) return fail("pmmc_tree_create: out is null", PMMC_ERR_INVALID);
  *out = nullptr;
  if (n_chr < 0 || n_regions < 0 || (n_chr > 0 && !chrs) || (n_regions > 0 && !regions))
    return fail("pmmc_tree_create: bad chromosome or region list", PMMC_ERR_INVALID);
  pmmc_tree *t = new pmmc_tree();
  for (int32_t r = 0; r < n_regions; ++r) {
    if (!regions[r].chr) {
      delete t;
      return fail("pmmc_tree_create: region without chromosome name", PMMC_ERR_INVALID);
    }
    t->regions.push_back(TRegion{regions[r].chr, regions[r].start, regions[r].end});
  }
  for (int32_t ci = 0; ci < n_chr; ++ci) {
    const pmmc_chromosome &in = chrs[ci];
    TChr c;
    c.name = in.name ? in.name : "";
    if (in.n_anchors < 0 || in.n_arcs < 0 || in.n_long_arcs < 0 ||
        (in.n_anchors > 0 && (!in.anchor_start || !in.anchor_end)) ||
        (in.n_arcs > 0 && (!in.arc_start || !in.arc_end || !in.arc_score)) ||
        (in.n_long_arcs > 0 && (!in.long_start || !in.long_end || !in.long_score))) {
      delete t;
      return fail("pmmc_tree_create: chromosome " + c.name + ": bad array", PMMC_ERR_INVALID);
    }
    if (in.n_anchors == 0) {
      delete t;
      return fail("pmmc_tree_create: chromosome " + c.name + " has no anchors", PMMC_ERR_INVALID);
    }
    for (int32_t i = 0; i < in.n_anchors; ++i) {
      c.anchors.push_back({in.anchor_start[i], in.anchor_end[i]});
      if (in.orientation) c.orient.push_back(in.orientation[i]);
    }
    for (int32_t i = 0; i < in.n_arcs; ++i) {
      if (in.arc_start[i] < 0 || in.arc_start[i] >= in.n_anchors || in.arc_end[i] < 0 ||
          in.arc_end[i] >= in.n_anchors) {
        delete t;
        return fail("pmmc_tree_create: chromosome " + c.name + ": arc endpoint out of range",
                    PMMC_ERR_INVALID);
      }
      c.arcs.push_back({in.arc_start[i], in.arc_end[i]});
      c.scores.push_back(in.arc_score[i]);
      // InteractionArc(start, end, score) (InteractionArc.hpp:103-120)
      c.arc_info.push_back({-1, -1, in.arc_score[i], -1});
    }
    for (int32_t i = 0; i < in.n_long_arcs; ++i)
      c.long_arcs.push_back({in.long_start[i], in.long_end[i], in.long_score[i]});
    t->chrs.push_back(std::move(c));
    if (!build_chromosome(*t, t->chrs.back())) {
      std::string m = g_err;
      delete t;
      return fail(m, PMMC_ERR_INVALID);
    }
  }
  *out = t;
  return PMMC_OK;
}

extern "C" void pmmc_tree_destroy(pmmc_tree *t) { delete t; }

extern "C" int32_t pmmc_tree_size(const pmmc_tree *t) { return t ? (int32_t)t->cl.size() : 0; }

extern "C" int32_t pmmc_tree_n_chromosomes(const pmmc_tree *t) {
  return t ? (int32_t)t->chrs.size() : 0;
}

namespace pmmc_host {

bool valid_cluster(const pmmc_tree *t, int32_t c) {
  return t && c >= 0 && c < (int32_t)t->cl.size();
}

pmmc_status copy_list(const std::vector<int32_t> &v, int32_t *out, int32_t cap, int32_t *n) {
  if (n) *n = (int32_t)v.size();
  if (cap > 0 && !out) return fail("output list is null", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < cap && i < (int32_t)v.size(); ++i) out[i] = v[i];
  return PMMC_OK;
}

// LooperSolver::setLevel = setTopLevel + levelDown (LooperSolver.hpp:2431-2452,
// :3087-3104): children in stored order
std::vector<int32_t> level_list(const pmmc_tree &t, int32_t chr, int32_t level) {
  std::vector<int32_t> cur(1, t.chrs[chr].root), nxt;
  for (int32_t l = 0; l < level; ++l) {
    nxt.clear();
    for (int32_t c : cur) {
      const std::vector<int32_t> &ch = t.cl[c].children;
      if (ch.empty())
        nxt.push_back(c);
      else
        nxt.insert(nxt.end(), ch.begin(), ch.end());
    }
    cur.swap(nxt);
  }
  return cur;
}

}  // namespace pmmc_host

extern "C" pmmc_status pmmc_tree_get(const pmmc_tree *t, int32_t first, int32_t count,
                                     pmmc_cluster *out) {
  if (!t || !out || first < 0 || count < 0 || first + count > (int32_t)t->cl.size())
    return fail("pmmc_tree_get: range outside the tree", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < count; ++i) {
    const TCluster &c = t->cl[first + i];
    pmmc_cluster &o = out[i];
    o.start = c.start;
    o.end = c.end;
    o.genomic_pos = c.genomic_pos;
    o.base_start = c.base_start;
    o.base_end = c.base_end;
    o.level = c.level;
    o.parent = c.parent;
    o.n_children = (int32_t)c.children.size();
    o.is_fixed = c.is_fixed ? 1 : 0;
    o.orientation = c.orientation;
    o.pos[0] = c.pos[0];
    o.pos[1] = c.pos[1];
    o.pos[2] = c.pos[2];
    o.dist_to_next = c.dist_to_next;
  }
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_tree_children(const pmmc_tree *t, int32_t cluster, int32_t *out,
                                          int32_t cap, int32_t *n) {
  if (!valid_cluster(t, cluster)) return fail("pmmc_tree_children: bad cluster", PMMC_ERR_INVALID);
  return copy_list(t->cl[cluster].children, out, cap, n);
}

extern "C" pmmc_status pmmc_tree_chromosome(const pmmc_tree *t, int32_t chr, int32_t *first,
                                            int32_t *root) {
  if (!t || chr < 0 || chr >= (int32_t)t->chrs.size())
    return fail("pmmc_tree_chromosome: bad chromosome", PMMC_ERR_INVALID);
  if (first) *first = t->chrs[chr].first;
  if (root) *root = t->chrs[chr].root;
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_tree_arcs(const pmmc_tree *t, int32_t chr, int32_t *out, int32_t cap,
                                      int32_t *n) {
  if (!t || chr < 0 || chr >= (int32_t)t->chrs.size())
    return fail("pmmc_tree_arcs: bad chromosome", PMMC_ERR_INVALID);
  const TChr &c = t->chrs[chr];
  if (n) *n = (int32_t)c.arcs.size();
  if (cap > 0 && !out) return fail("pmmc_tree_arcs: out is null", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < cap && i < (int32_t)c.arcs.size(); ++i) {
    out[2 * i] = c.arcs[i][0];
    out[2 * i + 1] = c.arcs[i][1];
  }
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_tree_level(const pmmc_tree *t, int32_t chr, int32_t level,
                                       int32_t *out, int32_t cap, int32_t *n) {
  if (!t || chr < 0 || chr >= (int32_t)t->chrs.size() || level < 0)
    return fail("pmmc_tree_level: bad chromosome or level", PMMC_ERR_INVALID);
  return copy_list(level_list(*t, chr, level), out, cap, n);
}

extern "C" pmmc_status pmmc_tree_set_positions(pmmc_tree *t, int32_t n, const int32_t *ids,
                                               const float *xyz) {
  if (!t || n < 0 || (n > 0 && (!ids || !xyz)))
    return fail("pmmc_tree_set_positions: bad arguments", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < n; ++i)
    if (!valid_cluster(t, ids[i])) return fail("pmmc_tree_set_positions: bad cluster", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) t->cl[ids[i]].pos[k] = xyz[3 * i + k];
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_tree_get_positions(const pmmc_tree *t, int32_t n, const int32_t *ids,
                                               float *xyz) {
  if (!t || n < 0 || (n > 0 && (!ids || !xyz)))
    return fail("pmmc_tree_get_positions: bad arguments", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < n; ++i)
    if (!valid_cluster(t, ids[i])) return fail("pmmc_tree_get_positions: bad cluster", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) xyz[3 * i + k] = t->cl[ids[i]].pos[k];
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_tree_set_fixed(pmmc_tree *t, int32_t n, const int32_t *ids,
                                           const int32_t *fixed) {
  if (!t || n < 0 || (n > 0 && (!ids || !fixed)))
    return fail("pmmc_tree_set_fixed: bad arguments", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < n; ++i)
    if (!valid_cluster(t, ids[i])) return fail("pmmc_tree_set_fixed: bad cluster", PMMC_ERR_INVALID);
  for (int32_t i = 0; i < n; ++i) t->cl[ids[i]].is_fixed = fixed[i] != 0;
  return PMMC_OK;
}

// ---------------------------------------------------------------- densify

namespace pmmc_host {

// LooperSolver::densifyActiveRegion (LooperSolver.hpp:3836-3898); positions
// interpolate(a, b, st) = a + (b - a) * st (common.cpp:176), float per axis
void densify(pmmc_tree &t, int32_t cluster, int32_t add, bool fix) {
  const std::vector<int32_t> active = t.cl[cluster].children;
  if (active.empty()) return;
  std::vector<int32_t> out;
  const int32_t level = t.cl[active[0]].level + 1;
  for (size_t i = 0; i + 1 < active.size(); ++i) {
    if (fix) t.cl[active[i]].is_fixed = true;
    const int32_t range = t.cl[active[i + 1]].start - t.cl[active[i]].end;
    const int32_t d = range / (add + 1);
    int32_t p = t.cl[active[i]].end;
    out.push_back(active[i]);
    const float dst = 1.0f / static_cast<float>(add + 1);
    float st = dst;
    for (int32_t j = 0; j < add; ++j) {
      p += d;
      TCluster c = make_cluster(p, p);
      const float *a = t.cl[active[i]].pos;
      const float *b = t.cl[active[i + 1]].pos;
      for (int k = 0; k < 3; ++k) {
        const float diff = b[k] - a[k];
        const float step = diff * st;
        c.pos[k] = a[k] + step;
      }
      c.level = level;
      t.cl.push_back(c);
      out.push_back((int32_t)t.cl.size() - 1);
      st += dst;
    }
  }
  out.push_back(active.back());
  if (fix) t.cl[active.back()].is_fixed = true;
  t.cl[cluster].children = out;
}

// dist_to_next / noise_size of reconstructClusterArcsDistances
// (LooperSolver.hpp:4339-4354); returns the mean spacing (before the
// coefficient)
float set_dist_to_next(pmmc_tree &t, const std::vector<int32_t> &active,
                       const pmmc_level_params *p) {
  float noise = 0.0f;
  for (size_t i = 0; i + 1 < active.size(); ++i) {
    const int32_t d = std::abs(t.cl[active[i + 1]].genomic_pos - t.cl[active[i]].genomic_pos);
    const float dist = static_cast<float>(pmmc_genomic_length_to_distance(d, p));
    t.cl[active[i]].dist_to_next = dist;
    noise += dist;
  }
  noise /= static_cast<float>(active.size());
  return noise;
}

}  // namespace pmmc_host

extern "C" pmmc_status pmmc_tree_densify(pmmc_tree *t, int32_t cluster, int32_t loop_density,
                                         int32_t fix, int32_t *n_active) {
  if (!valid_cluster(t, cluster) || loop_density < 0)
    return fail("pmmc_tree_densify: bad cluster or loop density", PMMC_ERR_INVALID);
  densify(*t, cluster, loop_density, fix != 0);
  if (n_active) *n_active = (int32_t)t->cl[cluster].children.size();
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_tree_set_dist_to_next(pmmc_tree *t, int32_t cluster,
                                                  const pmmc_level_params *p, float *noise_anchor,
                                                  float *noise_subanchor) {
  if (!valid_cluster(t, cluster) || !p)
    return fail("pmmc_tree_set_dist_to_next: bad cluster", PMMC_ERR_INVALID);
  const std::vector<int32_t> active = t->cl[cluster].children;
  if (active.empty()) return fail("pmmc_tree_set_dist_to_next: cluster has no children", PMMC_ERR_INVALID);
  const float noise = set_dist_to_next(*t, active, p);
  if (noise_anchor) *noise_anchor = noise * p->noise_anchor;
  if (noise_subanchor) *noise_subanchor = noise * p->noise_subanchor;
  return PMMC_OK;
}

// ---------------------------------------------------------------- expected distances

namespace pmmc_host {

// first anchor containing x (Anchor::contains, Anchor.hpp:87: start <= x <=
// end); binary search when the anchors are sorted and disjoint, where it
// finds the same anchor
AnchorLookup::AnchorLookup(const TChr &chr) : c(chr) {
  const int n_anc = (int)c.anchors.size();
  for (int k = 0; k + 1 < n_anc && disjoint; ++k)
    disjoint = c.anchors[k][0] <= c.anchors[k][1] && c.anchors[k][1] < c.anchors[k + 1][0];
}

int AnchorLookup::operator()(int32_t x) const {
  const int n_anc = (int)c.anchors.size();
  if (disjoint) {
    int lo = 0, hi = n_anc;
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (c.anchors[mid][1] < x)
        lo = mid + 1;
      else
        hi = mid;
    }
    return (lo < n_anc && c.anchors[lo][0] <= x && x <= c.anchors[lo][1]) ? lo : -1;
  }
  for (int k = 0; k < n_anc; ++k)
    if (x >= c.anchors[k][0] && x <= c.anchors[k][1]) return k;
  return -1;
}

// LooperSolver::calcAnchorExpectedDistancesHeatmap (LooperSolver.hpp:5494-5627)
bool expected_distances(const pmmc_tree &t, int32_t chr, const std::vector<int32_t> &active,
                        const pmmc_level_params *p, const float *anchor_heatmap, float *m) {
  const int n = (int)active.size();
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) m[(size_t)i * n + j] = i == j ? 0.0f : -1.0f;
  // cluster -> slot; operator[] on a missing key inserts slot 0, which is
  // where an arc leaving the block lands (:5538-5539)
  std::map<int32_t, int32_t> c2a;
  for (int i = 0; i < n; ++i) c2a[active[i]] = i;
  const TChr &c = t.chrs[chr];
  for (int i = 0; i < n; ++i) {
    const int32_t ai = active[i];
    for (int32_t a : t.cl[ai].arcs) {
      const int32_t other = other_end(c, a, ai);
      if (other == -1) {
        g_err = "expected distances: arc without the cluster on it";
        return false;
      }
      if (other < ai) continue;
      const float ed = pmmc_freq_to_distance(c.scores[a], p);
      const int32_t x = c2a[ai];
      const int32_t y = c2a[other];
      m[(size_t)x * n + y] = ed;
      m[(size_t)y * n + x] = ed;
    }
  }
  if (!c.long_arcs.empty() && !c.anchors.empty()) {
    const int n_anc = (int)c.anchors.size();
    std::vector<int32_t> a2a(n_anc, -1);
    for (const auto &kv : c2a)
      if (kv.first >= c.first && kv.first < c.first + n_anc) a2a[kv.first - c.first] = kv.second;
    const AnchorLookup pos_to_anc(c);
    for (const auto &la : c.long_arcs) {
      const int ia = pos_to_anc(la[0]), ib = pos_to_anc(la[1]);
      if (ia < 0 || ib < 0 || ia == ib) continue;
      const int32_t x = a2a[ia], y = a2a[ib];
      if (x < 0 || y < 0) continue;
      if (m[(size_t)x * n + y] >= 0.0f) continue;
      const float ed = pmmc_freq_to_distance(la[2], p);
      m[(size_t)x * n + y] = ed;
      m[(size_t)y * n + x] = ed;
    }
  }
  if (p->use_anchor_heatmap && anchor_heatmap) {
    const float max_val = heat_max(anchor_heatmap, n);
    if (static_cast<double>(max_val) > 1e-6) {
      for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
          float &v = m[(size_t)i * n + j];
          if (v < 0.0f) continue;
          float s = (anchor_heatmap[(size_t)i * n + j] / max_val) * p->anchor_heatmap_influence;
          if (s > 1.0f) s = 1.0f;
          v *= (1.0f - s);
          m[(size_t)j * n + i] = v;
        }
    }
  }
  return true;
}

}  // namespace pmmc_host

extern "C" pmmc_status pmmc_tree_expected_distances(const pmmc_tree *t, int32_t chr,
                                                    int32_t cluster, const pmmc_level_params *p,
                                                    const float *anchor_heatmap, float *out) {
  if (!valid_cluster(t, cluster) || !p || !out || chr < 0 || chr >= (int32_t)t->chrs.size())
    return fail("pmmc_tree_expected_distances: bad arguments", PMMC_ERR_INVALID);
  if (cluster < t->chrs[chr].first || cluster > t->chrs[chr].root)
    return fail("pmmc_tree_expected_distances: cluster is not in the chromosome", PMMC_ERR_INVALID);
  if (p->use_anchor_heatmap && !anchor_heatmap)
    return fail("pmmc_tree_expected_distances: use_anchor_heatmap without a heatmap",
                PMMC_ERR_INVALID);
  if (!expected_distances(*t, chr, t->cl[cluster].children, p, anchor_heatmap, out))
    return fail(g_err, PMMC_ERR_INVALID);
  return PMMC_OK;
}
