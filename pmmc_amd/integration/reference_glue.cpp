// reference_glue.cpp -- drop-in replacement for the reference's four .cu
// translation units on the annealing path (ParallelMonteCarloHeatmap.cu,
// ParallelMonteCarloArcs.cu, ParallelMonteCarloSmooth.cu, ParallelMarkArcs.cu).
//
// It is compiled INSIDE the reference source tree with the reference's own
// headers (g++, no CUDA), and calls libpmmc_hip.so through include/pmmc_hip.h.
// Each member flattens the solver state exactly as the .cu host code did,
// calls the C ABI, writes the positions back and returns the value the
// reference returned.  Build it from the reference's tree with
//   g++ -std=c++17 -O2 -Isrc/include -Isrc/thirdparty -I<repo>/include -c reference_glue.cpp
// in place of the four .cu files, and link libpmmc_hip.so.
//
// Two knobs a build or a harness may set (the reference's own code calls
// neither; declared in reference_glue.h): the chains per call
// (pmmc_glue_set_chains, 0 = the library default) and the Philox segment of
// the next call (pmmc_glue_set_segment; each call takes the next one).
// With PMMC_GLUE_TRACE=<file> in the environment every annealing call
// appends its flattened inputs, the positions it wrote back into
// clusters[], its return value and total_mc_steps to that file.
#include <LooperSolver.hpp>

#include <ctime>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pmmc_hip.h"
#include "reference_glue.h"

namespace {

pmmc_ctx *g_ctx = nullptr;       // one context per process (device 0)
unsigned int g_seed = 0;         // setGpuSeed
uint32_t g_call = 0;             // distinct Philox segment per call
int32_t g_chains = 0;            // chains per call (0: library default)

// ---- optional call trace (PMMC_GLUE_TRACE) ----
FILE *trace_file() {
  static FILE *f = nullptr;
  static bool tried = false;
  if (!tried) {
    tried = true;
    const char *path = std::getenv("PMMC_GLUE_TRACE");
    if (path && *path) f = std::fopen(path, "a");
  }
  return f;
}

void trace_floats(FILE *f, const char *tag, const float *v, size_t n) {
  std::fprintf(f, "%s %zu", tag, n);
  for (size_t i = 0; i < n; ++i) std::fprintf(f, " %a", v[i]);
  std::fprintf(f, "\n");
}

void trace_ints(FILE *f, const char *tag, const int32_t *v, size_t n) {
  std::fprintf(f, "%s %zu", tag, n);
  for (size_t i = 0; i < n; ++i) std::fprintf(f, " %d", v[i]);
  std::fprintf(f, "\n");
}

// header and inputs of one call
void trace_call(int level, const pmmc_run &run, int n, float step, long long steps_before,
                const std::vector<float> &pos, const std::vector<uint8_t> &fixed) {
  FILE *f = trace_file();
  if (!f) return;
  std::fprintf(f, "@@call %d %llu %u %d %d %a %lld\n", level,
               static_cast<unsigned long long>(run.seed), run.segment, run.n_chains, n,
               static_cast<double>(step), steps_before);
  trace_floats(f, "pos", pos.data(), pos.size());
  std::vector<int32_t> fx(fixed.begin(), fixed.end());
  trace_ints(f, "fixed", fx.data(), fx.size());
}

// what the call left in clusters[] (xyz of the active region) and returned
void trace_result(const std::vector<Cluster> &cl, const std::vector<int> &active, float ret,
                  long long steps_after) {
  FILE *f = trace_file();
  if (!f) return;
  std::vector<float> out;
  for (int c : active) {
    out.push_back(cl[c].pos.x);
    out.push_back(cl[c].pos.y);
    out.push_back(cl[c].pos.z);
  }
  trace_floats(f, "out", out.data(), out.size());
  std::fprintf(f, "ret %a %lld\n", static_cast<double>(ret), steps_after);
  std::fflush(f);
}

pmmc_ctx *ctx_or_die() {
  if (!g_ctx) {
    const pmmc_status st = pmmc_ctx_create(0, &g_ctx);
    if (st != PMMC_OK) {
      std::fprintf(stderr, "pmmc: no usable gfx950 device (status %d)\n", static_cast<int>(st));
      std::exit(1);
    }
  }
  return g_ctx;
}

void check(pmmc_status st, const char *what) {
  if (st != PMMC_OK) {  // the reference exit()s on any GPU error (gpuErrchk)
    std::fprintf(stderr, "pmmc: %s failed: %s\n", what, pmmc_last_error(g_ctx));
    std::exit(static_cast<int>(st));
  }
}

uint64_t run_seed() {
  return g_seed != 0 ? g_seed : static_cast<uint64_t>(std::time(nullptr));
}

pmmc_level_settings level_settings(double T0, double dt, double thr, float tjs, float tjc,
                                   int min_succ, int steps) {
  pmmc_level_settings s;
  s.max_temp = T0;
  s.dt_temp = dt;
  s.improvement = thr;
  s.tjs = tjs;
  s.tjc = tjc;
  s.min_successes = min_succ;
  s.steps = steps;
  s.milestone_fails = Settings::milestoneFailsThreshold;
  s.sphere_radius = Settings::sphere_radius;
  s.use2d = Settings::use2D ? 1 : 0;
  return s;
}

}  // namespace

// platform.h:106, ParallelMonteCarloHeatmap.cu:37
void setGpuSeed(unsigned int seed) {
  g_seed = seed;
  if (g_ctx) pmmc_set_gpu_seed(g_ctx, seed);
}

void pmmc_glue_set_chains(int32_t chains) { g_chains = chains > 0 ? chains : 0; }
void pmmc_glue_set_segment(uint32_t segment) { g_call = segment; }

// GpuResourceCache (LooperSolver.hpp:54-106): device memory now lives in the
// pmmc_ctx arena; the cache only records that the device is ready.
void GpuResourceCache::init() {
  if (initialized) return;
  sm_count = pmmc_ctx_cu_count(ctx_or_die());
  initialized = true;
}
void GpuResourceCache::initArcsBuffers(int, int, int) {}
void GpuResourceCache::initSmoothBuffers(int) {}
void GpuResourceCache::cleanup() { initialized = false; }

// ParallelMonteCarloHeatmap.cu:365-511
float LooperSolver::ParallelMonteCarloHeatmap(float step_size) {
  const int n = static_cast<int>(active_region.size());
  if (n <= 1) return 0.0f;
  std::vector<float> pos(3 * n), dist(static_cast<size_t>(n) * n);
  std::vector<uint8_t> fixed(n);
  for (int i = 0; i < n; ++i) {
    const Cluster &c = clusters[active_region[i]];
    pos[3 * i] = c.pos.x;
    pos[3 * i + 1] = c.pos.y;
    pos[3 * i + 2] = c.pos.z;
    fixed[i] = c.is_fixed ? 1 : 0;
  }
  const int hs = static_cast<int>(heatmap_dist.size);
  if (hs != n) return 0.0f;  // the kernel reports a size mismatch and scores 0
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) dist[static_cast<size_t>(i) * n + j] = heatmap_dist.v[i][j];
  pmmc_heatmap_problem p;
  p.n = n;
  p.pos = pos.data();
  p.fixed = fixed.data();
  p.dist = dist.data();
  p.diagonal = heatmap_dist.diagonal_size;
  p.boundaries = heatmap_chromosome_boundaries.empty() ? nullptr : heatmap_chromosome_boundaries.data();
  p.n_boundaries = static_cast<int32_t>(heatmap_chromosome_boundaries.size());
  p.skip_below = 1e-3f;  // ParallelMonteCarloHeatmap.cu:176
  const pmmc_level_settings ls = level_settings(
      Settings::maxTempHeatmap, Settings::dtTempHeatmap, Settings::MCstopConditionImprovementHeatmap,
      Settings::tempJumpScaleHeatmap, Settings::tempJumpCoefHeatmap,
      Settings::MCstopConditionMinSuccessesHeatmap, Settings::MCstopConditionStepsHeatmap);
  pmmc_schedule s;
  check(pmmc_schedule_preset(PMMC_LEVEL_HEATMAP, PMMC_PRESET_GPU, &ls, step_size, &s), "preset");
  const pmmc_run run = {run_seed(), g_call++, 0, g_chains};
  if (FILE *f = trace_file()) {
    trace_call(1, run, n, step_size, static_cast<long long>(total_mc_steps), pos, fixed);
    trace_floats(f, "mat", dist.data(), dist.size());
    std::vector<int32_t> bnd(heatmap_chromosome_boundaries.begin(), heatmap_chromosome_boundaries.end());
    std::fprintf(f, "diag %d\n", p.diagonal);
    trace_ints(f, "bnd", bnd.data(), bnd.size());
  }
  pmmc_anneal_stats st;
  check(pmmc_heatmap_anneal(ctx_or_die(), &p, &s, &run, pos.data(), &st), "heatmap anneal");
  for (int i = 0; i < n; ++i) {
    Cluster &c = clusters[active_region[i]];
    c.pos.x = pos[3 * i];
    c.pos.y = pos[3 * i + 1];
    c.pos.z = pos[3 * i + 2];
  }
  total_mc_steps += static_cast<int>(st.best_chain_moves);
  const float ret = static_cast<float>(st.energy_resummed);  // calcScoreHeatmapActiveRegion()
  trace_result(clusters, active_region, ret, static_cast<long long>(total_mc_steps));
  return ret;
}

// ParallelMonteCarloArcs.cu:680-944
float LooperSolver::parallelMonteCarloArcs(float step_size) {
  const int n = static_cast<int>(active_region.size());
  if (n <= 1) return 0.0f;
  rebuildActiveLoopMap();
  std::vector<float> pos(3 * n), ed(static_cast<size_t>(n) * n, -1.0f);
  std::vector<uint8_t> fixed(n);
  for (int i = 0; i < n; ++i) {
    const Cluster &c = clusters[active_region[i]];
    pos[3 * i] = c.pos.x;
    pos[3 * i + 1] = c.pos.y;
    pos[3 * i + 2] = c.pos.z;
    fixed[i] = c.is_fixed ? 1 : 0;
  }
  if (static_cast<int>(heatmap_exp_dist_anchor.size) == n)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) ed[static_cast<size_t>(i) * n + j] = heatmap_exp_dist_anchor.v[i][j];
  std::vector<int32_t> lp;
  std::vector<float> lq;
  for (const auto &e : active_loop_map) {
    lp.push_back(e.first.first);
    lp.push_back(e.first.second);
    lq.push_back(loop_constraints[e.second].stiffness);
    lq.push_back(loop_constraints[e.second].eq_distance);
  }
  pmmc_arcs_problem p;
  p.n = n;
  p.pos = pos.data();
  p.fixed = fixed.data();
  p.exp_dist = ed.data();
  p.n_loops = static_cast<int32_t>(lq.size() / 2);
  p.loop_pairs = lp.empty() ? nullptr : lp.data();
  p.loop_params = lq.empty() ? nullptr : lq.data();
  p.k_stretch = Settings::springConstantStretchArcs;
  p.k_squeeze = Settings::springConstantSqueezeArcs;
  const pmmc_level_settings ls = level_settings(
      Settings::maxTemp, Settings::dtTemp, Settings::MCstopConditionImprovement,
      Settings::tempJumpScale, Settings::tempJumpCoef, Settings::MCstopConditionMinSuccesses,
      Settings::MCstopConditionSteps);
  pmmc_schedule s;
  check(pmmc_schedule_preset(PMMC_LEVEL_ARCS, PMMC_PRESET_GPU, &ls, step_size, &s), "preset");
  const pmmc_run run = {run_seed(), g_call++, 0, g_chains};
  if (FILE *f = trace_file()) {
    trace_call(2, run, n, step_size, static_cast<long long>(total_mc_steps), pos, fixed);
    trace_floats(f, "mat", ed.data(), ed.size());
    trace_ints(f, "loops", lp.data(), lp.size());
    trace_floats(f, "loopk", lq.data(), lq.size());
  }
  pmmc_anneal_stats st;
  check(pmmc_arcs_anneal(ctx_or_die(), &p, &s, &run, pos.data(), &st), "arcs anneal");
  for (int i = 0; i < n; ++i) {
    Cluster &c = clusters[active_region[i]];
    c.pos.x = pos[3 * i];
    c.pos.y = pos[3 * i + 1];
    c.pos.z = pos[3 * i + 2];
  }
  total_mc_steps += static_cast<int>(st.best_chain_moves);
  if (sim_logger)
    sim_logger->logAcceptanceRatio("arcs_gpu", static_cast<int>(st.best_chain_moves),
                                   static_cast<int>(st.best_chain_accepted));
  const float ret = static_cast<float>(st.energy);  // the best chain's running score
  trace_result(clusters, active_region, ret, static_cast<long long>(total_mc_steps));
  return ret;
}

// ParallelMonteCarloSmooth.cu:357-506
float LooperSolver::ParallelMonteCarloSmooth(float step_size, unsigned long long seed) {
  const int n = static_cast<int>(active_region.size());
  if (n <= 2) return 0.0f;
  std::vector<float> pos(3 * n), dtn(n);
  std::vector<uint8_t> fixed(n);
  for (int i = 0; i < n; ++i) {
    const Cluster &c = clusters[active_region[i]];
    pos[3 * i] = c.pos.x;
    pos[3 * i + 1] = c.pos.y;
    pos[3 * i + 2] = c.pos.z;
    fixed[i] = c.is_fixed ? 1 : 0;
    dtn[i] = static_cast<float>(c.dist_to_next);
  }
  pmmc_smooth_problem p;
  p.n = n;
  p.pos = pos.data();
  p.fixed = fixed.data();
  p.dist_to_next = dtn.data();
  p.k_stretch = Settings::springConstantStretch;
  p.k_squeeze = Settings::springConstantSqueeze;
  p.k_angular = Settings::springAngularConstant;
  p.w_dist = Settings::weightDistSmooth;
  p.w_angle = Settings::weightAngleSmooth;
  const pmmc_level_settings ls = level_settings(
      Settings::maxTempSmooth, Settings::dtTempSmooth, Settings::MCstopConditionImprovementSmooth,
      Settings::tempJumpScaleSmooth, Settings::tempJumpCoefSmooth,
      Settings::MCstopConditionMinSuccessesSmooth, Settings::MCstopConditionStepsSmooth);
  pmmc_schedule s;
  check(pmmc_schedule_preset(PMMC_LEVEL_SMOOTH, PMMC_PRESET_GPU, &ls, step_size, &s), "preset");
  const pmmc_run run = {seed != 0 ? seed : run_seed(), g_call++, 0, g_chains};
  if (FILE *f = trace_file()) {
    trace_call(3, run, n, step_size, static_cast<long long>(total_mc_steps), pos, fixed);
    trace_floats(f, "mat", dtn.data(), dtn.size());
  }
  pmmc_anneal_stats st;
  check(pmmc_smooth_anneal(ctx_or_die(), &p, &s, &run, pos.data(), &st), "smooth anneal");
  for (int i = 0; i < n; ++i) {
    if (fixed[i]) continue;  // only non-fixed beads are written back
    Cluster &c = clusters[active_region[i]];
    c.pos.x = pos[3 * i];
    c.pos.y = pos[3 * i + 1];
    c.pos.z = pos[3 * i + 2];
  }
  total_mc_steps += static_cast<int>(st.best_chain_moves);
  if (sim_logger)
    sim_logger->logAcceptanceRatio("smooth_gpu", static_cast<int>(st.best_chain_moves),
                                   static_cast<int>(st.best_chain_accepted));
  const float ret = static_cast<float>(st.energy_resummed);  // calcScoreStructureSmooth(true, true)
  trace_result(clusters, active_region, ret, static_cast<long long>(total_mc_steps));
  return ret;
}

// ParallelMarkArcs.cu:69-244: all chromosomes in one call, grouping state
// (last_start and the unordered_map) carried across them as the reference does.
void InteractionArcs::parallelMarkArcs(bool /*ignore_missing*/) {
  const int C = static_cast<int>(chrs.size());
  std::vector<int32_t> anc_off(C + 1, 0), raw_off(C + 1, 0), as, ae, raw;
  for (int c = 0; c < C; ++c) {
    const std::string &chr = chrs[c];
    std::printf(" %s...\n", chr.c_str());
    std::sort(raw_arcs[chr].begin(), raw_arcs[chr].end());
    for (int i = 0; i < anchors_cnt[chr]; ++i) {
      as.push_back(anchors[chr][i].start);
      ae.push_back(anchors[chr][i].end);
    }
    for (const InteractionArc &a : raw_arcs[chr]) {
      raw.push_back(a.start);
      raw.push_back(a.end);
      raw.push_back(a.score);
      raw.push_back(a.factor);
    }
    anc_off[c + 1] = static_cast<int32_t>(as.size());
    raw_off[c + 1] = static_cast<int32_t>(raw.size() / 4);
  }
  std::vector<pmmc_arc> out(2 * raw.size() / 4 + 1);
  std::vector<int32_t> out_off(C + 1), nonmatching(C + 1);
  check(pmmc_group_arcs(ctx_or_die(), C, anc_off.data(), as.data(), ae.data(), raw_off.data(),
                        raw.data(), out.data(), out_off.data(), nonmatching.data()),
        "mark arcs");
  for (int c = 0; c < C; ++c) {
    const std::string &chr = chrs[c];
    arcs[chr].clear();
    for (int k = out_off[c]; k < out_off[c + 1]; ++k) {
      InteractionArc a(out[k].start, out[k].end, out[k].score, out[k].factor);
      a.genomic_start = out[k].genomic_start;
      a.genomic_end = out[k].genomic_end;
      a.eff_score = out[k].eff_score;
      arcs[chr].push_back(a);
    }
    if (nonmatching[c] > 0) std::printf("  %d non-matching arcs skipped\n", nonmatching[c]);
    arcs_cnt[chr] = static_cast<int>(arcs[chr].size());
  }
}
