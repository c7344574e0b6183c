// group.cpp -- oracle restatement of the host half of arc marking.
// TEST INFRASTRUCTURE ONLY.
//
// After every raw arc has been mapped to anchor indices (orc_mark_arcs), the
// reference groups arcs by start anchor, merges arcs between the same anchor
// pair per factor and emits a summary arc when several factors meet:
//   - InteractionArcs::markArcs        InteractionArcs.hpp:207-336 (std::map)
//   - InteractionArcs::parallelMarkArcs ParallelMarkArcs.cu:150-236
//     (std::unordered_map, whose libstdc++ iteration order decides the
//     emitted order; the map object and last_start live across chromosomes,
//     ParallelMarkArcs.cu:77-79).
// The raw arcs are sorted first with InteractionArc::operator<
// (InteractionArc.hpp:74-84: start, end, factor).
//
// PINNED BY: the std::map variant is compared arc-for-arc with the
// reference's own markArcs built into oracle/_ref (tests/test_oracle.py).
// The unordered_map variant emits the same multiset of arcs per start anchor
// (checked there too); its order is the libstdc++ unordered_map order.
#include <algorithm>
#include <cstdint>
#include <map>
#include <unordered_map>
#include <vector>

#include "oracle.h"

namespace {

struct Arc {
  int32_t start, end, gstart, gend, score, eff, factor;
  bool operator<(const Arc &o) const {
    if (start != o.start) return start < o.start;
    if (end != o.end) return end < o.end;
    return factor < o.factor;
  }
};

Arc make_arc(int s, int e, int score, int factor) {
  Arc a{s, e, -1, -1, score, score, factor};
  return a;
}

void emit_group(int end_anchor, std::vector<Arc> v, std::vector<Arc> &out) {
  if (v.size() == 1) {
    out.push_back(v[0]);
    return;
  }
  std::sort(v.begin(), v.end());
  bool multi = false;
  for (size_t j = 1; j < v.size(); ++j)
    if (v[j].factor != v[j - 1].factor) multi = true;
  int total = 0, fscore = 0;
  size_t first = 0;
  for (size_t j = 0; j <= v.size(); ++j) {
    if (j == v.size() || (j > 0 && v[j].factor != v[j - 1].factor)) {
      v[first].score = fscore;
      v[first].eff = multi ? 0 : fscore;
      out.push_back(v[first]);
      first = j;
      total += fscore;
      fscore = 0;
    }
    if (j < v.size()) fscore += v[j].score;
  }
  if (multi) {
    Arc s = make_arc(v[0].start, end_anchor, 0, -1);
    s.eff = total;
    out.push_back(s);
  }
}

template <class MapT>
int group_all(int n_chr, const int32_t *anc_off, const int32_t *anc_s, const int32_t *anc_e,
              const int32_t *arc_off, const int32_t *raw, int32_t *out, int32_t *out_off,
              int32_t *nonmatch) {
  MapT tmp;
  int last_start = -1;
  int total_out = 0;
  out_off[0] = 0;
  for (int c = 0; c < n_chr; ++c) {
    int A = anc_off[c + 1] - anc_off[c];
    int R = arc_off[c + 1] - arc_off[c];
    std::vector<Arc> rs;
    rs.reserve(R);
    for (int i = 0; i < R; ++i) {
      const int32_t *r = raw + 4 * (size_t)(arc_off[c] + i);
      rs.push_back(make_arc(r[0], r[1], r[2], r[3]));
    }
    std::sort(rs.begin(), rs.end());
    std::vector<int32_t> ss(R), ee(R), ms(R), me(R);
    for (int i = 0; i < R; ++i) { ss[i] = rs[i].start; ee[i] = rs[i].end; }
    orc_mark_arcs(anc_s + anc_off[c], anc_e + anc_off[c], A, ss.data(), ee.data(), R, ms.data(),
                  me.data());
    std::vector<Arc> res;
    int nm = 0;
    for (int i = 0; i <= R; ++i) {
      int st = -1, en = -1;
      if (i < R) {
        st = ms[i];
        en = me[i];
        if (st == -1 || en == -1) { ++nm; continue; }
        if (st == en) continue;
      }
      if (st != last_start || i == R) {
        for (auto &el : tmp) emit_group(el.first, el.second, res);
        tmp.clear();
        last_start = st;
      }
      if (i < R) {
        Arc a = make_arc(st, en, rs[i].score, rs[i].factor);
        a.gstart = rs[i].start;
        a.gend = rs[i].end;
        tmp[en].push_back(a);
      }
    }
    tmp.clear();
    if (nonmatch) nonmatch[c] = nm;
    for (const Arc &a : res) {
      int32_t *o = out + 7 * (size_t)total_out;
      o[0] = a.start; o[1] = a.end; o[2] = a.gstart; o[3] = a.gend;
      o[4] = a.score; o[5] = a.eff; o[6] = a.factor;
      ++total_out;
    }
    out_off[c + 1] = total_out;
  }
  return total_out;
}

}  // namespace

extern "C" {

// raw: [R_total*4] (genomic start, genomic end, score, factor) per chromosome
// slice arc_off; out: [R_total*2*7] is always large enough.  Returns the
// number of emitted arcs.  use_unordered: 1 = parallelMarkArcs order,
// 0 = markArcs order.
int orc_group_arcs(int use_unordered, int n_chr, const int32_t *anc_off, const int32_t *anc_s,
                   const int32_t *anc_e, const int32_t *arc_off, const int32_t *raw,
                   int32_t *out, int32_t *out_off, int32_t *nonmatch) {
  if (use_unordered)
    return group_all<std::unordered_map<int, std::vector<Arc>>>(n_chr, anc_off, anc_s, anc_e,
                                                                arc_off, raw, out, out_off,
                                                                nonmatch);
  return group_all<std::map<int, std::vector<Arc>>>(n_chr, anc_off, anc_s, anc_e, arc_off, raw,
                                                    out, out_off, nonmatch);
}

}  // extern "C"
