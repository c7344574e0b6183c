// markarcs.hip -- InteractionArcs::parallelMarkArcs (ParallelMarkArcs.cu:69-244).
//
// Device part (ParallelMarkArcs.cu:36-64): for every raw arc endpoint find
// the LAST anchor j with length_j > 1 that contains it.  When a chromosome's
// anchor starts are sorted (the reference loads them sorted), the candidates
// for position x are the prefix [0, k) with k = #(start_j <= x), and the
// answer is the rightmost j < k whose effective end (end_j if length_j > 1,
// else -inf) is >= x: one binary search plus one query on a max-segment tree,
// O(log A) per endpoint instead of the reference's O(A) scan.  Unsorted
// chromosomes fall back to the full scan, which has the same result.
//
// Host part (ParallelMarkArcs.cu:150-243): grouping by start anchor and
// per-factor merging, reproducing the reference's std::unordered_map
// emission order and its cross-chromosome last_start.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "ctx.hpp"

namespace pmmc {

struct MarkChr {
  int32_t anc0, A;        // anchors slice
  int32_t tree0, S;       // max-tree slice (2*S nodes), S = 0 -> unsorted
  int32_t arc0, R;        // arcs slice
};

__device__ __forceinline__ bool anchor_ok(int s, int e) {
  const int len = (s == 0 && e == 0) ? 0 : e - s + 1;  // Anchor::length, Anchor.hpp:81-85
  return len > 1;
}

__device__ int mark_tree(const int32_t *__restrict__ starts, const int32_t *__restrict__ tree, int A,
                         int S, int x) {
  // k = number of anchors with start <= x (upper bound)
  int lo = 0, hi = A;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (starts[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  int h = S + lo, l = S;
  while (l < h) {
    if (h & 1) {
      --h;
      if (tree[h] >= x) {
        while (h < S) {
          h = 2 * h + 1;
          if (tree[h] < x) --h;
        }
        return h - S;
      }
    }
    l >>= 1;
    h >>= 1;
  }
  return -1;
}

__device__ int mark_scan(const int32_t *__restrict__ s, const int32_t *__restrict__ e, int A, int x) {
  int r = -1;
  for (int j = 0; j < A; ++j)
    if (anchor_ok(s[j], e[j]) && x >= s[j] && x <= e[j]) r = j;
  return r;
}

__global__ __launch_bounds__(256) void k_mark(const MarkChr *__restrict__ chrs, int n_chr,
                                              const int32_t *__restrict__ anc_s,
                                              const int32_t *__restrict__ anc_e,
                                              const int32_t *__restrict__ tree,
                                              const int32_t *__restrict__ arc_s,
                                              const int32_t *__restrict__ arc_e, int total,
                                              int32_t *__restrict__ out_s, int32_t *__restrict__ out_e) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    int lo = 0, hi = n_chr - 1;  // chromosome of arc i
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (chrs[mid].arc0 <= i) lo = mid;
      else hi = mid - 1;
    }
    const MarkChr c = chrs[lo];
    const int32_t *s = anc_s + c.anc0, *e = anc_e + c.anc0;
    if (c.S > 0) {
      const int32_t *t = tree + c.tree0;
      out_s[i] = mark_tree(s, t, c.A, c.S, arc_s[i]);
      out_e[i] = mark_tree(s, t, c.A, c.S, arc_e[i]);
    } else {
      out_s[i] = mark_scan(s, e, c.A, arc_s[i]);
      out_e[i] = mark_scan(s, e, c.A, arc_e[i]);
    }
  }
}

// Build the per-chromosome max trees on the host and run the kernel.
static pmmc_status mark_all(pmmc_ctx *ctx, int n_chr, const int32_t *anc_off, const int32_t *anc_s,
                            const int32_t *anc_e, const int32_t *arc_off, const int32_t *arc_s,
                            const int32_t *arc_e, int32_t *out_s, int32_t *out_e) {
  const int A_tot = anc_off[n_chr], R_tot = arc_off[n_chr];
  if (R_tot == 0) return PMMC_OK;
  std::vector<MarkChr> chrs(n_chr);
  std::vector<int32_t> tree;
  for (int c = 0; c < n_chr; ++c) {
    MarkChr &m = chrs[c];
    m.anc0 = anc_off[c];
    m.A = anc_off[c + 1] - anc_off[c];
    m.arc0 = arc_off[c];
    m.R = arc_off[c + 1] - arc_off[c];
    m.tree0 = static_cast<int32_t>(tree.size());
    m.S = 0;
    bool sorted = true;
    for (int j = 1; j < m.A; ++j)
      if (anc_s[m.anc0 + j] < anc_s[m.anc0 + j - 1]) sorted = false;
    if (sorted && m.A > 0) {
      int S = 1;
      while (S < m.A) S <<= 1;
      m.S = S;
      tree.resize(tree.size() + 2 * S, INT_MIN);
      int32_t *t = tree.data() + m.tree0;
      for (int j = 0; j < m.A; ++j) {
        const int s = anc_s[m.anc0 + j], e = anc_e[m.anc0 + j];
        const int len = (s == 0 && e == 0) ? 0 : e - s + 1;
        t[S + j] = len > 1 ? e : INT_MIN;
      }
      for (int k = S - 1; k >= 1; --k) t[k] = std::max(t[2 * k], t[2 * k + 1]);
    }
  }
  if (tree.empty()) tree.push_back(INT_MIN);
  Sizer z;
  z.add<MarkChr>(n_chr);
  z.add<int32_t>(std::max(A_tot, 1));
  z.add<int32_t>(std::max(A_tot, 1));
  z.add<int32_t>(tree.size());
  for (int k = 0; k < 4; ++k) z.add<int32_t>(R_tot);
  pmmc_status ps = arena_reserve(ctx, z.top + 4096);
  if (ps != PMMC_OK) return ps;
  Carver cv{ctx->arena, 0, ctx->arena_cap};
  MarkChr *d_chrs = cv.take<MarkChr>(n_chr);
  int32_t *d_as = cv.take<int32_t>(std::max(A_tot, 1));
  int32_t *d_ae = cv.take<int32_t>(std::max(A_tot, 1));
  int32_t *d_tree = cv.take<int32_t>(tree.size());
  int32_t *d_rs = cv.take<int32_t>(R_tot);
  int32_t *d_re = cv.take<int32_t>(R_tot);
  int32_t *d_os = cv.take<int32_t>(R_tot);
  int32_t *d_oe = cv.take<int32_t>(R_tot);
  hipStream_t s = ctx->stream;
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_chrs, chrs.data(), sizeof(MarkChr) * n_chr, hipMemcpyHostToDevice, s));
  if (A_tot > 0) {
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_as, anc_s, sizeof(int32_t) * A_tot, hipMemcpyHostToDevice, s));
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_ae, anc_e, sizeof(int32_t) * A_tot, hipMemcpyHostToDevice, s));
  }
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_tree, tree.data(), sizeof(int32_t) * tree.size(), hipMemcpyHostToDevice, s));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_rs, arc_s, sizeof(int32_t) * R_tot, hipMemcpyHostToDevice, s));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_re, arc_e, sizeof(int32_t) * R_tot, hipMemcpyHostToDevice, s));
  const int grid = std::min((R_tot + 255) / 256, 4096);
  hipLaunchKernelGGL(k_mark, dim3(grid), dim3(256), 0, s, d_chrs, n_chr, d_as, d_ae, d_tree, d_rs, d_re,
                     R_tot, d_os, d_oe);
  PMMC_HIP_TRY(ctx, hipGetLastError());
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(out_s, d_os, sizeof(int32_t) * R_tot, hipMemcpyDeviceToHost, s));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(out_e, d_oe, sizeof(int32_t) * R_tot, hipMemcpyDeviceToHost, s));
  PMMC_HIP_TRY(ctx, hipStreamSynchronize(s));
  return PMMC_OK;
}

// ---------------- host grouping ----------------
struct Arc {
  int32_t start, end, gstart, gend, score, eff, factor;
};

// InteractionArc::operator< (InteractionArc.hpp:74-84)
static bool arc_less(const Arc &a, const Arc &b) {
  if (a.start != b.start) return a.start < b.start;
  if (a.end != b.end) return a.end < b.end;
  return a.factor < b.factor;
}

// One group of arcs sharing (start anchor, end anchor): ParallelMarkArcs.cu:176-229
static void emit(int end_anchor, std::vector<Arc> v, std::vector<Arc> &out) {
  if (v.size() == 1) {
    out.push_back(v[0]);
    return;
  }
  std::sort(v.begin(), v.end(), arc_less);
  bool multi = false;
  for (size_t j = 1; j < v.size(); ++j) multi |= v[j].factor != v[j - 1].factor;
  int total = 0, fscore = 0;
  size_t first = 0;
  for (size_t j = 0; j <= v.size(); ++j) {
    if (j == v.size() || (j > 0 && v[j].factor != v[j - 1].factor)) {
      Arc a = v[first];
      a.score = fscore;
      a.eff = multi ? 0 : fscore;
      out.push_back(a);
      first = j;
      total += fscore;
      fscore = 0;
    }
    if (j < v.size()) fscore += v[j].score;
  }
  if (multi) out.push_back(Arc{v[0].start, end_anchor, -1, -1, 0, total, -1});
}

}  // namespace pmmc

using namespace pmmc;

extern "C" pmmc_status pmmc_mark_arcs(pmmc_ctx *ctx, const int32_t *anchor_start,
                                      const int32_t *anchor_end, int32_t n_anchors,
                                      const int32_t *arc_start, const int32_t *arc_end, int32_t n_arcs,
                                      int32_t *out_start, int32_t *out_end) {
  if (!ctx || n_anchors < 0 || n_arcs < 0) return fail(ctx, PMMC_ERR_INVALID, "mark_arcs: bad sizes");
  if ((n_anchors > 0 && (!anchor_start || !anchor_end)) ||
      (n_arcs > 0 && (!arc_start || !arc_end || !out_start || !out_end)))
    return fail(ctx, PMMC_ERR_INVALID, "mark_arcs: null array");
  const int32_t ao[2] = {0, n_anchors}, ro[2] = {0, n_arcs};
  return mark_all(ctx, 1, ao, anchor_start, anchor_end, ro, arc_start, arc_end, out_start, out_end);
}

extern "C" pmmc_status pmmc_group_arcs(pmmc_ctx *ctx, int32_t n_chr, const int32_t *anc_off,
                                       const int32_t *anc_start, const int32_t *anc_end,
                                       const int32_t *raw_off, const int32_t *raw, pmmc_arc *out,
                                       int32_t *out_off, int32_t *nonmatching) {
  if (!ctx || n_chr < 0 || !anc_off || !raw_off || !out_off) return fail(ctx, PMMC_ERR_INVALID, "group_arcs: null");
  for (int c = 0; c < n_chr; ++c)
    if (anc_off[c + 1] < anc_off[c] || raw_off[c + 1] < raw_off[c])
      return fail(ctx, PMMC_ERR_INVALID, "group_arcs: offsets not monotone");
  const int R_tot = raw_off[n_chr];
  // 1. sort raw arcs of every chromosome (ParallelMarkArcs.cu:96)
  std::vector<Arc> sorted(R_tot);
  for (int i = 0; i < R_tot; ++i)
    sorted[i] = Arc{raw[4 * i], raw[4 * i + 1], -1, -1, raw[4 * i + 2], raw[4 * i + 2], raw[4 * i + 3]};
  for (int c = 0; c < n_chr; ++c)
    std::sort(sorted.begin() + raw_off[c], sorted.begin() + raw_off[c + 1], arc_less);
  // 2. mark every endpoint on the device
  std::vector<int32_t> ss(R_tot), ee(R_tot), ms(R_tot), me(R_tot);
  for (int i = 0; i < R_tot; ++i) {
    ss[i] = sorted[i].start;
    ee[i] = sorted[i].end;
  }
  pmmc_status ps = mark_all(ctx, n_chr, anc_off, anc_start, anc_end, raw_off, ss.data(), ee.data(),
                            ms.data(), me.data());
  if (ps != PMMC_OK) return ps;
  // 3. group by start anchor, merge per factor; the map and last_start
  //    persist across chromosomes exactly as in ParallelMarkArcs.cu:77-79
  std::unordered_map<int, std::vector<Arc>> tmp;
  int last_start = -1;
  int n_out = 0;
  out_off[0] = 0;
  std::vector<Arc> res;
  for (int c = 0; c < n_chr; ++c) {
    res.clear();
    int nm = 0;
    const int R0 = raw_off[c], R = raw_off[c + 1] - raw_off[c];
    for (int i = 0; i <= R; ++i) {
      int st = -1, en = -1;
      if (i < R) {
        st = ms[R0 + i];
        en = me[R0 + i];
        if (st == -1 || en == -1) {
          ++nm;
          continue;
        }
        if (st == en) continue;
      }
      if (st != last_start || i == R) {
        for (auto &el : tmp) emit(el.first, el.second, res);
        tmp.clear();
        last_start = st;
      }
      if (i < R) {
        const Arc &r = sorted[R0 + i];
        tmp[en].push_back(Arc{st, en, r.start, r.end, r.score, r.score, r.factor});
      }
    }
    tmp.clear();
    if (nonmatching) nonmatching[c] = nm;
    for (const Arc &a : res)
      out[n_out++] = pmmc_arc{a.start, a.end, a.gstart, a.gend, a.score, a.eff, a.factor};
    out_off[c + 1] = n_out;
  }
  return PMMC_OK;
}
