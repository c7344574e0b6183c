// Level drivers over the cluster tree (include/pmmc_levels.h): the heatmap
// level (reconstructClustersHeatmapSingleLevel, LooperSolver.hpp:1258-1422)
// and the arc / sub-anchor levels (reconstructClustersArcsDistances' block
// loop :3978-4140 over reconstructClusterArcsDistances :4303-4487).
//
// The reference loops over blocks and restarts on the host, one annealing
// launch each.  Here every restart of every block of a stage is one job of a
// single pmmc_anneal_batch launch; the host only prepares the jobs, picks the
// best restart per block and densifies between the two arc stages.
//
// Plain host C++ (g++, -ffp-contract=off): the restart noise is computed in
// float with separate roundings, as the reference's vector3 arithmetic.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <unordered_set>
#include <vector>

#include "levels_host.hpp"

using namespace pmmc_host;

namespace {

// Philox4x32-10 (Salmon et al., SC'11), the same stream the kernels draw
uint32_t mulhi(uint32_t a, uint32_t b, uint32_t *lo) {
  const uint64_t p = static_cast<uint64_t>(a) * b;
  *lo = static_cast<uint32_t>(p);
  return static_cast<uint32_t>(p >> 32);
}

void philox(uint32_t c[4], uint64_t seed) {
  uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
  for (int r = 0; r < 10; ++r) {
    uint32_t lo0, lo1;
    const uint32_t hi0 = mulhi(0xD2511F53u, c[0], &lo0);
    const uint32_t hi1 = mulhi(0xCD9E8D57u, c[2], &lo1);
    const uint32_t n0 = hi1 ^ c[1] ^ k0;
    const uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[1] = lo1;
    c[3] = lo0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

constexpr uint32_t kLevelInit = 5;

// random_vector(amplitude, use2d) (common.cpp:14-25) on the Philox stream:
// (2u - 1) * amplitude per axis, u = 24-bit uniform in [0, 1)
void restart_noise(uint64_t seed, uint32_t segment, uint32_t bead, float amp, bool use2d,
                   float out[3]) {
  uint32_t c[4] = {bead, 0u, segment, kLevelInit << 28};
  philox(c, seed);
  for (int k = 0; k < 3; ++k) {
    const float u = static_cast<float>(c[k] >> 8) * 0x1p-24f;
    const float t = 2.0f * u - 1.0f;
    out[k] = t * amp;
  }
  if (use2d) out[2] = 0.0f;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// One job of a stage: its own copy of every array its problem points to.
struct StageJob {
  int32_t level = 0;
  int32_t block = 0;             // index into the stage's block list
  std::vector<int32_t> beads;    // tree clusters of the problem, in order
  std::vector<float> pos, pos_out, mat, dtn;
  std::vector<int32_t> bnd;
  std::vector<uint8_t> fixed;
  pmmc_heatmap_problem hp{};
  pmmc_arcs_problem ap{};
  pmmc_smooth_problem sp{};
  pmmc_anneal_stats st{};
  pmmc_job job{};
};

// chains per job: the library's default pool for the stage's largest
// problem (pmmc_default_chains: whole rounds of the kernel's resident
// wavefronts, at least 16 per compute unit, positions within a third of the
// device's HBM) shared among the stage's jobs, at least one wavefront each
int32_t chains_for(pmmc_ctx *ctx, int32_t requested, size_t n_jobs, int32_t level, int32_t max_n) {
  if (requested > 0) return requested;
  const int32_t pool = pmmc_default_chains(ctx, level, max_n > 0 ? max_n : 1, 1);
  size_t ens = (static_cast<size_t>(pool > 0 ? pool : 64) / 64) / (n_jobs ? n_jobs : 1);
  if (ens < 1) ens = 1;
  return static_cast<int32_t>(ens) * 64;
}

// runs the stage's jobs as one batch, or (one_call) each as its own plain
// call in order, as the reference dispatches them; the results are the same
// bit for bit.  Adds moves / accepted / jobs to the report.
pmmc_status run_stage(pmmc_ctx *ctx, std::vector<std::unique_ptr<StageJob>> &jobs,
                      pmmc_level_report *rep, bool one_call) {
  if (jobs.empty()) return PMMC_OK;
  pmmc_status st = PMMC_OK;
  if (one_call) {
    for (auto &j : jobs) {
      const pmmc_job &b = j->job;
      if (b.level == PMMC_LEVEL_HEATMAP)
        st = pmmc_heatmap_anneal(ctx, static_cast<const pmmc_heatmap_problem *>(b.problem),
                                 &b.schedule, &b.run, b.pos_out, b.stats);
      else if (b.level == PMMC_LEVEL_ARCS)
        st = pmmc_arcs_anneal(ctx, static_cast<const pmmc_arcs_problem *>(b.problem), &b.schedule,
                              &b.run, b.pos_out, b.stats);
      else
        st = pmmc_smooth_anneal(ctx, static_cast<const pmmc_smooth_problem *>(b.problem),
                                &b.schedule, &b.run, b.pos_out, b.stats);
      if (st != PMMC_OK) break;
    }
  } else {
    std::vector<pmmc_job> js;
    js.reserve(jobs.size());
    for (auto &j : jobs) js.push_back(j->job);
    pmmc_batch_opts bo{};
    st = pmmc_anneal_batch(ctx, js.data(), (int32_t)js.size(), &bo);
  }
  if (st != PMMC_OK) return fail(std::string("level driver: ") + pmmc_last_error(ctx), st);
  for (auto &j : jobs) {
    rep->moves += j->st.moves;
    rep->accepted += j->st.accepted;
  }
  rep->n_jobs += (int32_t)jobs.size();
  return PMMC_OK;
}

// the reference's best-of rule: score < best || best < 0 (first on ties), on
// the value the annealing member returns: a float from the GPU members
// (LooperSolver.hpp:1392-1402, :4457-4467: `double score = <float member>()`),
// a double from the CPU annealer (MonteCarloArcsSmooth)
struct Best {
  double score = -1.0;
  int job = -1;
  void offer(double s, int j, bool gpu_member = true) {
    const double f = gpu_member ? static_cast<double>(static_cast<float>(s)) : s;
    if (f < score || score < 0.0) {
      score = f;
      job = j;
    }
  }
};


// stressRefineAnchors' problem over the active region `active` of chromosome
// `chr` (LooperSolver.hpp:5642-5760): the region sorted by genomic position,
// the pair list (arcs, then long arcs; first pair of two anchors kept) and
// the spring targets
struct StressInput {
  int32_t chr = -1;
  std::vector<int32_t> active;
  std::vector<float> pos, bb, pt;
  std::vector<int32_t> pi, pj;
};

void stress_input(const pmmc_tree &t, int32_t chr, std::vector<int32_t> active,
                  const pmmc_level_params *p, float ks, float kq, StressInput &in) {
  // std::sort with the reference's comparator, so equal genomic positions
  // (if any) end in the same order as there
  std::sort(active.begin(), active.end(),
            [&](int a, int b) { return t.cl[a].genomic_pos < t.cl[b].genomic_pos; });
  const TChr &c = t.chrs[chr];
  const int N = static_cast<int>(active.size());
  const int n_anc = static_cast<int>(c.anchors.size());
  const int cstart = c.first;
  std::vector<int> anc_to_active(n_anc, -1);
  for (int i = 0; i < N; ++i)
    if (active[i] >= cstart && active[i] < cstart + n_anc) anc_to_active[active[i] - cstart] = i;
  // beads adjacent in the sorted order are left to the springs while either
  // spring constant is on
  const bool backbone_active = ks > 1e-6f || kq > 1e-6f;
  std::unordered_set<int64_t> seen;
  auto add_pair = [&](int ai, int aj, float tgt) {
    if (ai == aj) return;
    if (ai > aj) std::swap(ai, aj);
    if (backbone_active && aj - ai <= 1) return;
    if (!seen.insert(static_cast<int64_t>(ai) * N + aj).second) return;
    in.pi.push_back(ai);
    in.pj.push_back(aj);
    in.pt.push_back(tgt);
  };
  for (size_t k = 0; k < c.arcs.size(); ++k) {
    const int a = c.arcs[k][0], b = c.arcs[k][1];
    if (a < cstart || a >= cstart + n_anc || b < cstart || b >= cstart + n_anc) continue;
    const int ai = anc_to_active[a - cstart], aj = anc_to_active[b - cstart];
    if (ai < 0 || aj < 0) continue;
    const float tt = pmmc_freq_to_distance(c.scores[k], p);
    if (tt > 1e-6f) add_pair(ai, aj, tt);
  }
  const AnchorLookup look(c);
  for (const auto &la : c.long_arcs) {
    const int ia = look(la[0]), ib = look(la[1]);
    if (ia < 0 || ib < 0) continue;
    const int as = anc_to_active[ia], bs = anc_to_active[ib];
    if (as < 0 || bs < 0) continue;
    const float tt = pmmc_freq_to_distance(la[2], p);
    if (tt > 1e-6f) add_pair(as, bs, tt);
  }
  in.bb.assign(N > 0 ? N : 1, -1.0f);
  for (int i = 0; i + 1 < N; ++i) {
    const int gap = std::abs(t.cl[active[i + 1]].genomic_pos - t.cl[active[i]].genomic_pos);
    if (gap > 0) in.bb[i] = static_cast<float>(pmmc_genomic_length_to_distance(gap, p));
  }
  in.pos.resize(3 * static_cast<size_t>(N));
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < 3; ++k) in.pos[3 * i + k] = t.cl[active[i]].pos[k];
  in.chr = chr;
  in.active = std::move(active);
}

// one launch over every input; refined positions written back to the tree
pmmc_status run_stress(pmmc_ctx *ctx, pmmc_tree &t, std::vector<StressInput> &ins, float ks,
                       float kq, float ka, const pmmc_stress_opts *opts,
                       std::vector<pmmc_stress_stats> &stats) {
  std::vector<pmmc_stress_problem> probs(ins.size());
  std::vector<float *> outs(ins.size());
  for (size_t i = 0; i < ins.size(); ++i) {
    StressInput &in = ins[i];
    probs[i] = pmmc_stress_problem{static_cast<int32_t>(in.active.size()), in.pos.data(), in.bb.data(),
                                   static_cast<int32_t>(in.pi.size()), in.pi.data(), in.pj.data(),
                                   in.pt.data(), ks, kq, ka};
    outs[i] = in.pos.data();
  }
  stats.assign(ins.size(), pmmc_stress_stats{});
  const pmmc_status st = pmmc_stress_refine(ctx, static_cast<int32_t>(ins.size()), probs.data(), opts,
                                            outs.data(), stats.data());
  if (st != PMMC_OK) return fail(std::string("stress refinement: ") + pmmc_last_error(ctx), st);
  for (StressInput &in : ins)
    for (size_t i = 0; i < in.active.size(); ++i)
      for (int k = 0; k < 3; ++k) t.cl[in.active[i]].pos[k] = in.pos[3 * i + k];
  return PMMC_OK;
}

// the chromosome's anchors, the active region of the global refinement
// (LooperSolver.hpp:4158-4162)
std::vector<int32_t> chromosome_anchors(const pmmc_tree &t, int32_t c) {
  std::vector<int32_t> a(t.chrs[c].anchors.size());
  for (size_t i = 0; i < a.size(); ++i) a[i] = t.chrs[c].first + static_cast<int32_t>(i);
  return a;
}

// Sub-anchor beads re-placed between the refined anchors
// (LooperSolver.hpp:4179-4239): every bead of the anchors' level + 1 among
// the first `limit` clusters whose genomic position lies within
// [first anchor, last anchor] of `active` (in its order) takes the position
// on the line between the anchors around it.  Returns the beads moved.
int32_t reinterpolate(pmmc_tree &t, const std::vector<int32_t> &active, size_t limit) {
  if (active.empty()) return 0;
  const int32_t sub_level = t.cl[active[0]].level + 1;
  std::vector<int32_t> gp(active.size());
  for (size_t i = 0; i < active.size(); ++i) gp[i] = t.cl[active[i]].genomic_pos;
  int32_t moved = 0;
  for (size_t ci = 0; ci < limit && ci < t.cl.size(); ++ci) {
    TCluster &b = t.cl[ci];
    if (b.level != sub_level) continue;
    const int32_t g = b.genomic_pos;
    if (g < gp.front() || g > gp.back()) continue;
    const size_t j = static_cast<size_t>(std::lower_bound(gp.begin(), gp.end(), g) - gp.begin());
    if (j == 0) {
      for (int k = 0; k < 3; ++k) b.pos[k] = t.cl[active[0]].pos[k];
    } else if (j >= gp.size()) {
      for (int k = 0; k < 3; ++k) b.pos[k] = t.cl[active.back()].pos[k];
    } else {
      const int32_t g0 = gp[j - 1], g1 = gp[j];
      const float u = g1 > g0 ? static_cast<float>(g - g0) / static_cast<float>(g1 - g0) : 0.5f;
      const float *p0 = t.cl[active[j - 1]].pos, *p1 = t.cl[active[j]].pos;
      for (int k = 0; k < 3; ++k) b.pos[k] = (1.0f - u) * p0[k] + u * p1[k];
    }
    ++moved;
  }
  return moved;
}

}  // namespace

extern "C" pmmc_status pmmc_tree_stress_refine(pmmc_ctx *ctx, pmmc_tree *t, int32_t chr,
                                               const pmmc_level_params *p, float k_stretch,
                                               float k_squeeze, float k_angular,
                                               const pmmc_stress_opts *opts,
                                               pmmc_stress_stats *stats) {
  if (!ctx || !t || !p || !opts) return fail("pmmc_tree_stress_refine: null argument", PMMC_ERR_INVALID);
  if (chr < -1 || chr >= (int32_t)t->chrs.size())
    return fail("pmmc_tree_stress_refine: bad chromosome", PMMC_ERR_INVALID);
  std::vector<StressInput> ins;
  for (int32_t c = 0; c < (int32_t)t->chrs.size(); ++c) {
    if (chr >= 0 && c != chr) continue;
    ins.emplace_back();
    stress_input(*t, c, chromosome_anchors(*t, c), p, k_stretch, k_squeeze, ins.back());
  }
  std::vector<pmmc_stress_stats> st;
  const pmmc_status ps = run_stress(ctx, *t, ins, k_stretch, k_squeeze, k_angular, opts, st);
  if (ps != PMMC_OK) return ps;
  if (stats)
    for (size_t i = 0; i < st.size(); ++i) stats[i] = st[i];
  return PMMC_OK;
}

namespace {

// Global anchor refinement of the listed chromosomes (LooperSolver.hpp:
// 4141-4243): one batched stage of every chromosome's anchor restarts, one
// stress launch, then the re-placement of sub-anchor beads chromosome by
// chromosome, in the reference's order.
pmmc_status global_stage(pmmc_ctx *ctx, pmmc_tree &t, const std::vector<int32_t> &chrs,
                         const pmmc_level_params *p, const pmmc_arcs_level_opts *o, uint64_t seed,
                         const std::vector<uint32_t> &gseg, const std::vector<size_t> &limit,
                         pmmc_level_report &rep) {
  const double t0 = now_ms();
  std::vector<int32_t> todo;
  for (int32_t c : chrs)
    if (t.chrs[c].anchors.size() > 1) todo.push_back(c);
  std::vector<std::unique_ptr<StageJob>> jobs;
  int32_t max_n = 1;
  for (int32_t c : todo) max_n = std::max<int32_t>(max_n, (int32_t)t.chrs[c].anchors.size());
  const size_t n_jobs = todo.size() * (size_t)std::max(0, p->restarts_anchor);
  const int32_t chains = chains_for(ctx, o->chains_per_restart, n_jobs, PMMC_LEVEL_ARCS, max_n);
  for (size_t gi = 0; gi < todo.size(); ++gi) {
    const int32_t c = todo[gi];
    const std::vector<int32_t> active = chromosome_anchors(t, c);
    const int n = (int)active.size();
    std::vector<float> ed((size_t)n * n);
    if (!expected_distances(t, c, active, p, nullptr, ed.data())) return fail(g_err, PMMC_ERR_INVALID);
    const float noise_size = set_dist_to_next(t, active, p) * p->noise_anchor;
    pmmc_schedule sched{};
    pmmc_status st = pmmc_schedule_preset(PMMC_LEVEL_ARCS, o->arcs_preset, &o->arcs, noise_size, &sched);
    if (st != PMMC_OK) return fail("pmmc_arcs_level: arc schedule preset", st);
    for (int32_t k = 0; k < p->restarts_anchor; ++k) {
      auto j = std::make_unique<StageJob>();
      const uint32_t jseg = gseg[c] + (uint32_t)k;
      j->level = PMMC_LEVEL_ARCS;
      j->block = (int32_t)gi;
      j->beads = active;
      j->pos.resize((size_t)n * 3);
      j->fixed.resize(n);
      for (int i = 0; i < n; ++i) {
        // restart noise on every bead at the anchor level (:4415-4420)
        const TCluster &b = t.cl[active[i]];
        float d[3];
        restart_noise(seed, jseg, (uint32_t)i, p->noise_anchor_small, p->use2d != 0, d);
        for (int a = 0; a < 3; ++a) j->pos[3 * i + a] = b.pos[a] + d[a];
        j->fixed[i] = b.is_fixed ? 1 : 0;
      }
      j->pos_out = j->pos;
      j->mat = ed;
      j->ap = pmmc_arcs_problem{n, j->pos.data(), j->fixed.data(), j->mat.data(), 0, nullptr,
                                nullptr, o->k_stretch_arcs, o->k_squeeze_arcs};
      j->job.level = PMMC_LEVEL_ARCS;
      j->job.problem = &j->ap;
      j->job.schedule = sched;
      j->job.run = pmmc_run{seed, jseg, 0u, chains};
      j->job.pos_out = j->pos_out.data();
      j->job.stats = &j->st;
      jobs.push_back(std::move(j));
    }
  }
  pmmc_status st = run_stage(ctx, jobs, &rep, o->one_call_per_job != 0);
  if (st != PMMC_OK) return st;
  {
    std::vector<Best> best(todo.size());
    for (size_t ji = 0; ji < jobs.size(); ++ji) best[jobs[ji]->block].offer(jobs[ji]->st.energy, (int)ji);
    for (size_t gi = 0; gi < todo.size(); ++gi) {
      if (best[gi].job < 0) continue;
      const StageJob &j = *jobs[best[gi].job];
      for (size_t i = 0; i < j.beads.size(); ++i)
        for (int a = 0; a < 3; ++a) t.cl[j.beads[i]].pos[a] = j.pos_out[3 * i + a];
      rep.score_global += best[gi].score;
    }
  }
  jobs.clear();
  // stress descent over the anchors sorted by genomic position (:4175-4178)
  std::vector<std::vector<int32_t>> order(todo.size());
  for (size_t gi = 0; gi < todo.size(); ++gi) order[gi] = chromosome_anchors(t, todo[gi]);
  if (o->stress.max_iter > 0) {
    std::vector<StressInput> ins(todo.size());
    for (size_t gi = 0; gi < todo.size(); ++gi)
      stress_input(t, todo[gi], order[gi], p, o->k_stretch, o->k_squeeze, ins[gi]);
    std::vector<pmmc_stress_stats> sst;
    st = run_stress(ctx, t, ins, o->k_stretch, o->k_squeeze, o->k_angular, &o->stress, sst);
    if (st != PMMC_OK) return st;
    for (size_t gi = 0; gi < todo.size(); ++gi) {
      order[gi] = ins[gi].active;
      rep.stress_iterations += sst[gi].iterations;
    }
  }
  for (size_t gi = 0; gi < todo.size(); ++gi)
    rep.reinterpolated += reinterpolate(t, order[gi], limit[todo[gi]]);
  rep.ms_global = now_ms() - t0;
  return PMMC_OK;
}
}  // namespace

extern "C" void pmmc_arcs_level_opts_default(pmmc_arcs_level_opts *o) {
  std::memset(o, 0, sizeof(*o));
  // [simulation_arcs] / [simulation_arcs_smooth] / [cuda] defaults of
  // Settings::init (Settings.hpp:397, :478-494); the float settings keep
  // their float values
  o->arcs.max_temp = 20.0f;
  o->arcs.dt_temp = 0.99995f;
  o->arcs.improvement = 0.995f;
  o->arcs.tjs = 50.0f;
  o->arcs.tjc = 20.0f;
  o->arcs.min_successes = 5;
  o->arcs.steps = 10000;
  o->arcs.milestone_fails = 3;
  o->smooth = o->arcs;
  o->arcs_preset = PMMC_PRESET_GPU;
  o->smooth_preset = PMMC_PRESET_GPU;
  o->k_stretch_arcs = 1.0f;
  o->k_squeeze_arcs = 1.0f;
  o->k_stretch = 0.1f;
  o->k_squeeze = 0.1f;
  o->k_angular = 0.1f;
  o->w_dist = 1.0f;
  o->w_angle = 1.0f;
  o->chains_per_restart = 0;
  o->densify = 1;
  o->global_refine = 0;  // useGlobalAnchorRefinement (Settings.hpp:406)
  pmmc_stress_opts_default(&o->stress);
}

extern "C" void pmmc_heatmap_level_opts_default(pmmc_heatmap_level_opts *o) {
  std::memset(o, 0, sizeof(*o));
  // [simulation_heatmap] / [cuda] defaults (Settings.hpp:397, :462-468)
  o->heatmap.max_temp = 20.0f;
  o->heatmap.dt_temp = 0.99995f;
  o->heatmap.improvement = 0.995f;
  o->heatmap.tjs = 50.0f;
  o->heatmap.tjc = 20.0f;
  o->heatmap.min_successes = 5;
  o->heatmap.steps = 10000;
  o->heatmap.milestone_fails = 3;
  o->preset = PMMC_PRESET_GPU;
  o->chains_per_restart = 0;
}

// ---------------------------------------------------------------- arcs level

extern "C" pmmc_status pmmc_arcs_level(pmmc_ctx *ctx, pmmc_tree *t, int32_t chr,
                                       const pmmc_level_params *p, const pmmc_arcs_level_opts *o,
                                       uint64_t seed, uint32_t segment,
                                       pmmc_level_report *report) {
  if (!ctx || !t || !p || !o) return fail("pmmc_arcs_level: null argument", PMMC_ERR_INVALID);
  if (chr < -1 || chr >= (int32_t)t->chrs.size())
    return fail("pmmc_arcs_level: bad chromosome", PMMC_ERR_INVALID);
  if (p->restarts_anchor < 0 || p->restarts_subanchor < 0 || p->loop_density < 0)
    return fail("pmmc_arcs_level: negative restarts or loop density", PMMC_ERR_INVALID);
  pmmc_level_report rep{};

  struct Block {
    int32_t chr, cluster;
  };
  std::vector<Block> blocks;
  for (int32_t c = 0; c < (int32_t)t->chrs.size(); ++c) {
    if (chr >= 0 && c != chr) continue;
    for (int32_t ib : level_list(*t, c, 2))
      if (t->cl[ib].level == 3) blocks.push_back({c, ib});
  }
  rep.n_blocks = (int32_t)blocks.size();
  // Philox segments in the order the reference driver issues its GPU calls
  // (:4027-4138: per block, its anchor restarts, then densification and its
  // sub-anchor restarts; a block of at most one child makes no call, and a
  // sub-anchor problem of fewer than 4 beads runs the CPU annealer there,
  // :4431-4437), so a run matches the reference's driver over the
  // reference-side glue, whose per-call segment counter starts at `segment`.
  // Jobs that are not GPU calls there take segments from 2^30 on.
  // The global refinement's restarts follow each chromosome's blocks.
  const int32_t n_chr = (int32_t)t->chrs.size();
  std::vector<int32_t> chr_list;
  for (int32_t c = 0; c < n_chr; ++c)
    if (chr < 0 || c == chr) chr_list.push_back(c);
  std::vector<uint32_t> aseg(blocks.size(), 0), sseg(blocks.size(), 0), gseg(n_chr, 0);
  {
    uint32_t call = 0, cpu = 0;
    size_t bi = 0;
    for (int32_t c : chr_list) {
      for (; bi < blocks.size() && blocks[bi].chr == c; ++bi) {
        const size_t na = t->cl[blocks[bi].cluster].children.size();
        if (na <= 1) continue;
        aseg[bi] = segment + call;
        call += (uint32_t)p->restarts_anchor;
        const size_t ns = na + (na - 1) * (size_t)p->loop_density;  // after densification
        if (ns >= 4) {
          sseg[bi] = segment + call;
          call += (uint32_t)p->restarts_subanchor;
        } else {
          sseg[bi] = segment + (1u << 30) + cpu;
          cpu += (uint32_t)p->restarts_subanchor;
        }
      }
      if (o->global_refine && t->chrs[c].anchors.size() > 1) {
        gseg[c] = segment + call;
        call += (uint32_t)p->restarts_anchor;
      }
    }
  }
  // clusters that exist when each chromosome's global pass runs there: the
  // tree plus the sub-anchor beads of this and earlier chromosomes
  std::vector<size_t> limit(n_chr, t->cl.size());

  // ---- anchor stage: every restart of every block with >= 2 children
  std::vector<std::unique_ptr<StageJob>> jobs;
  int32_t max_n = 1;
  for (const Block &b : blocks) max_n = std::max<int32_t>(max_n, (int32_t)t->cl[b.cluster].children.size());
  size_t n_jobs_a = 0;
  for (const Block &b : blocks)
    if (t->cl[b.cluster].children.size() > 1) n_jobs_a += (size_t)p->restarts_anchor;
  const int32_t chains_a = chains_for(ctx, o->chains_per_restart, n_jobs_a, PMMC_LEVEL_ARCS, max_n);
  for (size_t bi = 0; bi < blocks.size(); ++bi) {
    const Block &b = blocks[bi];
    const std::vector<int32_t> active = t->cl[b.cluster].children;
    // children start at the block's position (LooperSolver.hpp:4051-4052)
    for (int32_t c : active)
      for (int k = 0; k < 3; ++k) t->cl[c].pos[k] = t->cl[b.cluster].pos[k];
    if (active.size() <= 1) continue;  // nothing to position (:4309-4310)
    const int n = (int)active.size();
    std::vector<float> ed((size_t)n * n);
    if (!expected_distances(*t, b.chr, active, p, nullptr, ed.data())) return fail(g_err, PMMC_ERR_INVALID);
    const float noise_size = set_dist_to_next(*t, active, p) * p->noise_anchor;
    pmmc_schedule sched{};
    pmmc_status st = pmmc_schedule_preset(PMMC_LEVEL_ARCS, o->arcs_preset, &o->arcs, noise_size, &sched);
    if (st != PMMC_OK) return fail("pmmc_arcs_level: arc schedule preset", st);
    for (int32_t k = 0; k < p->restarts_anchor; ++k) {
      auto j = std::make_unique<StageJob>();
      const uint32_t jseg = aseg[bi] + (uint32_t)k;
      j->level = PMMC_LEVEL_ARCS;
      j->block = (int32_t)bi;
      j->beads = active;
      j->pos.resize((size_t)n * 3);
      j->fixed.resize(n);
      for (int i = 0; i < n; ++i) {
        const TCluster &c = t->cl[active[i]];
        float d[3];
        restart_noise(seed, jseg, (uint32_t)i, p->noise_anchor_small, p->use2d != 0, d);
        for (int a = 0; a < 3; ++a) j->pos[3 * i + a] = c.pos[a] + d[a];
        j->fixed[i] = c.is_fixed ? 1 : 0;
      }
      j->pos_out = j->pos;
      j->mat = ed;
      j->ap = pmmc_arcs_problem{n, j->pos.data(), j->fixed.data(), j->mat.data(), 0, nullptr,
                                nullptr, o->k_stretch_arcs, o->k_squeeze_arcs};
      j->job.level = PMMC_LEVEL_ARCS;
      j->job.problem = &j->ap;
      j->job.schedule = sched;
      j->job.run = pmmc_run{seed, jseg, 0u, chains_a};
      j->job.pos_out = j->pos_out.data();
      j->job.stats = &j->st;
      jobs.push_back(std::move(j));
    }
  }
  double t0 = now_ms();
  pmmc_status st = run_stage(ctx, jobs, &rep, o->one_call_per_job != 0);
  if (st != PMMC_OK) return st;
  rep.ms_anchor = now_ms() - t0;
  {
    std::vector<Best> best(blocks.size());
    for (size_t ji = 0; ji < jobs.size(); ++ji) best[jobs[ji]->block].offer(jobs[ji]->st.energy, (int)ji);
    for (size_t bi = 0; bi < blocks.size(); ++bi) {
      if (best[bi].job < 0) continue;
      const StageJob &j = *jobs[best[bi].job];
      for (size_t i = 0; i < j.beads.size(); ++i)
        for (int a = 0; a < 3; ++a) t->cl[j.beads[i]].pos[a] = j.pos_out[3 * i + a];
      rep.score_anchor += best[bi].score;
    }
  }
  jobs.clear();
  if (!o->densify) {
    if (o->global_refine) {
      st = global_stage(ctx, *t, chr_list, p, o, seed, gseg, limit, rep);
      if (st != PMMC_OK) return st;
    }
    if (report) *report = rep;
    return PMMC_OK;
  }

  // ---- densify every block, then the sub-anchor stage
  {
    size_t bi = 0;
    for (int32_t c : chr_list) {
      for (; bi < blocks.size() && blocks[bi].chr == c; ++bi)
        densify(*t, blocks[bi].cluster, p->loop_density, true);
      limit[c] = t->cl.size();
    }
  }
  max_n = 1;
  size_t n_jobs_s = 0;
  for (const Block &b : blocks) {
    max_n = std::max<int32_t>(max_n, (int32_t)t->cl[b.cluster].children.size());
    if (t->cl[b.cluster].children.size() > 1) n_jobs_s += (size_t)p->restarts_subanchor;
  }
  const int32_t chains_s = chains_for(ctx, o->chains_per_restart, n_jobs_s, PMMC_LEVEL_SMOOTH, max_n);
  for (size_t bi = 0; bi < blocks.size(); ++bi) {
    const Block &b = blocks[bi];
    const std::vector<int32_t> active = t->cl[b.cluster].children;
    if (active.size() <= 1) continue;
    const int n = (int)active.size();
    const float noise_size = set_dist_to_next(*t, active, p) * p->noise_subanchor;
    // MonteCarloArcsSmooth below 4 beads, the GPU annealer otherwise (:4431-4437)
    const int preset = n >= 4 ? o->smooth_preset : PMMC_PRESET_CPU;
    pmmc_schedule sched{};
    st = pmmc_schedule_preset(PMMC_LEVEL_SMOOTH, preset, &o->smooth, noise_size, &sched);
    if (st != PMMC_OK) return fail("pmmc_arcs_level: smooth schedule preset", st);
    for (int32_t k = 0; k < p->restarts_subanchor; ++k) {
      auto j = std::make_unique<StageJob>();
      const uint32_t jseg = sseg[bi] + (uint32_t)k;
      j->level = PMMC_LEVEL_SMOOTH;
      j->block = (int32_t)bi;
      j->beads = active;
      j->pos.resize((size_t)n * 3);
      j->fixed.resize(n);
      j->dtn.resize(n);
      for (int i = 0; i < n; ++i) {
        const TCluster &c = t->cl[active[i]];
        float d[3] = {0.0f, 0.0f, 0.0f};
        if (!c.is_fixed) restart_noise(seed, jseg, (uint32_t)i, noise_size, p->use2d != 0, d);
        for (int a = 0; a < 3; ++a) j->pos[3 * i + a] = c.pos[a] + d[a];
        j->fixed[i] = c.is_fixed ? 1 : 0;
        j->dtn[i] = static_cast<float>(c.dist_to_next);
      }
      j->pos_out = j->pos;
      j->sp = pmmc_smooth_problem{n, j->pos.data(), j->fixed.data(), j->dtn.data(), o->k_stretch,
                                  o->k_squeeze, o->k_angular, o->w_dist, o->w_angle};
      j->job.level = PMMC_LEVEL_SMOOTH;
      j->job.problem = &j->sp;
      j->job.schedule = sched;
      j->job.run = pmmc_run{seed, jseg, 0u, chains_s};
      j->job.pos_out = j->pos_out.data();
      j->job.stats = &j->st;
      jobs.push_back(std::move(j));
    }
  }
  t0 = now_ms();
  st = run_stage(ctx, jobs, &rep, o->one_call_per_job != 0);
  if (st != PMMC_OK) return st;
  rep.ms_smooth = now_ms() - t0;
  {
    std::vector<Best> best(blocks.size());
    for (size_t ji = 0; ji < jobs.size(); ++ji)
      best[jobs[ji]->block].offer(jobs[ji]->st.energy_resummed, (int)ji,
                                  jobs[ji]->beads.size() >= 4);
    for (size_t bi = 0; bi < blocks.size(); ++bi) {
      if (best[bi].job < 0) continue;
      const StageJob &j = *jobs[best[bi].job];
      for (size_t i = 0; i < j.beads.size(); ++i)
        for (int a = 0; a < 3; ++a) t->cl[j.beads[i]].pos[a] = j.pos_out[3 * i + a];
      rep.score_smooth += best[bi].score;
    }
  }
  jobs.clear();
  if (o->global_refine) {
    st = global_stage(ctx, *t, chr_list, p, o, seed, gseg, limit, rep);
    if (st != PMMC_OK) return st;
  }
  if (report) *report = rep;
  return PMMC_OK;
}

// ---------------------------------------------------------------- heatmap level

extern "C" pmmc_status pmmc_heatmap_level(pmmc_ctx *ctx, pmmc_tree *t, int32_t level,
                                          const float *dist, int32_t diagonal,
                                          const pmmc_level_params *p,
                                          const pmmc_heatmap_level_opts *o, uint64_t seed,
                                          uint32_t segment, pmmc_level_report *report) {
  if (!ctx || !t || !p || !o || !dist) return fail("pmmc_heatmap_level: null argument", PMMC_ERR_INVALID);
  if (level != 0 && level != 1) return fail("pmmc_heatmap_level: level must be 0 or 1", PMMC_ERR_INVALID);
  pmmc_level_report rep{};
  const int32_t n_chr = (int32_t)t->chrs.size();
  std::vector<std::vector<int32_t>> cur(n_chr);
  for (int32_t c = 0; c < n_chr; ++c) cur[c] = level_list(*t, c, level);
  std::vector<int32_t> active;
  std::vector<int32_t> bnd;
  for (int32_t c = 0; c < n_chr; ++c) active.insert(active.end(), cur[c].begin(), cur[c].end());
  const int n = (int)active.size();
  if (n == 0) return fail("pmmc_heatmap_level: empty level", PMMC_ERR_INVALID);
  const float avg = heat_avg(dist, n);
  const float avg_dist = avg * (level == 0 ? p->noise_chr : p->noise_segment);
  if (level == 1) {
    // every segment starts at its chromosome's position (:1273-1290)
    for (int32_t c = 0; c < n_chr; ++c) {
      const int32_t par = t->cl[cur[c][0]].parent;
      if (par < 0) return fail("pmmc_heatmap_level: segment without a parent", PMMC_ERR_INVALID);
      float origin[3] = {t->cl[par].pos[0], t->cl[par].pos[1], t->cl[par].pos[2]};
      for (int32_t s : cur[c])
        for (int k = 0; k < 3; ++k) t->cl[s].pos[k] = origin[k];
    }
    // chromosome boundaries of the segment heatmap (:1221-1229)
    if (n_chr > 1) {
      int32_t acc = 0;
      bnd.push_back(acc);
      for (int32_t c = 0; c < n_chr; ++c) {
        acc += (int32_t)cur[c].size();
        bnd.push_back(acc);
      }
    }
  }
  const int32_t restarts = level == 0 ? p->restarts_chr : p->restarts_segment;
  if (restarts < 0) return fail("pmmc_heatmap_level: negative restarts", PMMC_ERR_INVALID);
  pmmc_schedule sched{};
  pmmc_status st = pmmc_schedule_preset(PMMC_LEVEL_HEATMAP, o->preset, &o->heatmap, avg_dist, &sched);
  if (st != PMMC_OK) return fail("pmmc_heatmap_level: schedule preset", st);
  const int32_t chains = chains_for(ctx, o->chains_per_restart, (size_t)restarts, PMMC_LEVEL_HEATMAP, n);
  std::vector<std::unique_ptr<StageJob>> jobs;
  for (int32_t k = 0; k < restarts; ++k) {
    auto j = std::make_unique<StageJob>();
    const uint32_t jseg = segment + (uint32_t)k;
    j->level = PMMC_LEVEL_HEATMAP;
    j->beads = active;
    j->pos.resize((size_t)n * 3);
    j->fixed.resize(n);
    for (int i = 0; i < n; ++i) {
      const TCluster &c = t->cl[active[i]];
      float d[3];
      restart_noise(seed, jseg, (uint32_t)i, avg_dist, p->use2d != 0, d);
      for (int a = 0; a < 3; ++a) j->pos[3 * i + a] = c.pos[a] + d[a];
      j->fixed[i] = c.is_fixed ? 1 : 0;
    }
    j->pos_out = j->pos;
    j->mat.assign(dist, dist + (size_t)n * n);
    j->bnd = bnd;
    // the GPU path skips targets below 1e-3, the CPU annealer below 1e-6
    // (ParallelMonteCarloHeatmap.cu:176, LooperSolver.hpp:3604)
    j->hp = pmmc_heatmap_problem{n, j->pos.data(), j->fixed.data(), j->mat.data(), diagonal,
                                 j->bnd.empty() ? nullptr : j->bnd.data(), (int32_t)j->bnd.size(),
                                 o->preset == PMMC_PRESET_GPU ? 1e-3f : 1e-6f};
    j->job.level = PMMC_LEVEL_HEATMAP;
    j->job.problem = &j->hp;
    j->job.schedule = sched;
    j->job.run = pmmc_run{seed, jseg, 0u, chains};
    j->job.pos_out = j->pos_out.data();
    j->job.stats = &j->st;
    jobs.push_back(std::move(j));
  }
  rep.n_blocks = 1;
  const double t0 = now_ms();
  st = run_stage(ctx, jobs, &rep, o->one_call_per_job != 0);
  if (st != PMMC_OK) return st;
  rep.ms_anchor = now_ms() - t0;
  Best best;
  for (size_t ji = 0; ji < jobs.size(); ++ji) best.offer(jobs[ji]->st.energy_resummed, (int)ji);
  if (best.job >= 0) {
    const StageJob &j = *jobs[best.job];
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) t->cl[active[i]].pos[a] = j.pos_out[3 * i + a];
    rep.score_anchor = best.score;
  }
  if (report) *report = rep;
  return PMMC_OK;
}
