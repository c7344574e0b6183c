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
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
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

// runs the stage's jobs as one batch; returns moves / accepted
pmmc_status run_stage(pmmc_ctx *ctx, std::vector<std::unique_ptr<StageJob>> &jobs,
                      pmmc_level_report *rep) {
  if (jobs.empty()) return PMMC_OK;
  std::vector<pmmc_job> js;
  js.reserve(jobs.size());
  for (auto &j : jobs) js.push_back(j->job);
  pmmc_batch_opts bo{};
  const pmmc_status st = pmmc_anneal_batch(ctx, js.data(), (int32_t)js.size(), &bo);
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
  std::vector<uint32_t> aseg(blocks.size(), 0), sseg(blocks.size(), 0);
  {
    uint32_t call = 0, cpu = 0;
    for (size_t bi = 0; bi < blocks.size(); ++bi) {
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
  }

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
  pmmc_status st = run_stage(ctx, jobs, &rep);
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
    if (report) *report = rep;
    return PMMC_OK;
  }

  // ---- densify every block, then the sub-anchor stage
  for (const Block &b : blocks) densify(*t, b.cluster, p->loop_density, true);
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
  st = run_stage(ctx, jobs, &rep);
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
  st = run_stage(ctx, jobs, &rep);
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
