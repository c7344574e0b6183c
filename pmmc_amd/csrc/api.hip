// api.hip -- context lifetime, schedule presets and problem preparation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "engine_host.hpp"

using namespace pmmc;

// ------------------------------------------------------------------ //
// context
// ------------------------------------------------------------------ //
extern "C" pmmc_status pmmc_ctx_create(int device, pmmc_ctx **out) {
  if (!out) return PMMC_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return PMMC_ERR_NODEVICE;
  if (device < 0 || device >= count) return PMMC_ERR_INVALID;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return PMMC_ERR_NODEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return PMMC_ERR_NODEVICE;
  if (hipSetDevice(device) != hipSuccess) return PMMC_ERR_NODEVICE;
  pmmc_ctx *c = new pmmc_ctx();
  c->device = device;
  c->cu_count = prop.multiProcessorCount;
  c->hbm_bytes = prop.totalGlobalMem;
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess &&
            hipEventCreate(&c->ev2) == hipSuccess;
  for (int k = 0; ok && k < kCtxSideStreams; ++k) {
    hipStream_t st = nullptr;
    ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
    if (ok) c->side.push_back(st);
  }
  if (!ok) {
    (void)hipGetLastError();
    pmmc_ctx_destroy(c);
    return PMMC_ERR_HIP;
  }
  // the allocation-free entry point's scratch (device current); a failure
  // only disables that entry point, which then reports it
  if (scratch_init(c) != PMMC_OK) c->scratch_error = c->last_error;
  *out = c;
  return PMMC_OK;
}

extern "C" void pmmc_ctx_destroy(pmmc_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->arena) hipFree(c->arena);
  if (c->host_stage) hipHostFree(c->host_stage);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->ev2) hipEventDestroy(c->ev2);
  for (hipStream_t st : c->side) {
    hipStreamSynchronize(st);
    hipStreamDestroy(st);
  }
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

extern "C" pmmc_status pmmc_ctx_reserve(pmmc_ctx *c, size_t bytes) {
  if (!c) return PMMC_ERR_INVALID;
  return arena_reserve(c, bytes);
}

extern "C" const char *pmmc_last_error(const pmmc_ctx *c) {
  return c ? c->last_error.c_str() : "null context";
}

extern "C" int pmmc_ctx_cu_count(const pmmc_ctx *c) { return c ? c->cu_count : 0; }

extern "C" void pmmc_set_gpu_seed(pmmc_ctx *c, uint64_t seed) {
  if (c) c->default_seed = seed;
}

extern "C" int32_t pmmc_pick_shift(int32_t level, int32_t n) { return pick_shift_of(level, n); }

extern "C" int32_t pmmc_default_chains(const pmmc_ctx *c, int32_t level, int32_t n,
                                       int32_t symmetric) {
  if (!c || n < 1 || level < PMMC_LEVEL_HEATMAP || level > PMMC_LEVEL_SMOOTH) return 0;
  HostProblem h;
  h.level = level;
  h.n = n;
  h.symmetric = symmetric != 0;
  // the row stride prep_heatmap gives the weight tables (selects the kernel)
  h.ldw = level == PMMC_LEVEL_HEATMAP ? (n <= 32 ? 32 : (n <= 64 ? 64 : n)) : 0;
  return default_chains(c, h);
}

// ------------------------------------------------------------------ //
// schedule presets: the reference annealers as pmmc_schedule
// ------------------------------------------------------------------ //
static pmmc_status fill_preset(int level, int preset, const pmmc_level_settings *s,
                               float step_size, pmmc_schedule *o);

extern "C" pmmc_status pmmc_schedule_preset(int level, int preset, const pmmc_level_settings *s,
                                            float step_size, pmmc_schedule *o) {
  if (!s || !o) return PMMC_ERR_INVALID;
  const pmmc_status st = fill_preset(level, preset, s, step_size, o);
  if (st != PMMC_OK) return st;
  // FP64 re-sum of the running total at most once per kResumMoves (8192)
  // moves: the GPU arc and heatmap kernels (256 / 512-move milestones)
  // re-sum at every 32nd / 16th milestone, the others (milestones every
  // `steps` moves) at every milestone.  The reference kernels keep a running
  // score only.
  o->resum = o->milestone_period >= kResumMoves ? 1 : (kResumMoves + o->milestone_period - 1) / o->milestone_period;
  return PMMC_OK;
}

static pmmc_status fill_preset(int level, int preset, const pmmc_level_settings *s,
                               float step_size, pmmc_schedule *o) {
  std::memset(o, 0, sizeof(*o));
  o->T0 = s->max_temp;
  o->dt = s->dt_temp;
  o->improvement = s->improvement;
  o->tjs = s->tjs;
  o->tjc = s->tjc;
  o->sphere_radius = s->sphere_radius;
  o->use2d = s->use2d;
  o->min_successes = s->min_successes;
  o->fails_threshold = s->milestone_fails;
  o->step_decay = 1.0f;
  o->resum = 1;
  o->guard_energy = -INFINITY;
  o->max_moves = INT32_MAX;
  const int steps = s->steps > 0 ? s->steps : 1;
  if (preset == PMMC_PRESET_CPU) {
    // milestone every `steps` counted moves, stop at the first failing one
    o->block = steps;
    o->milestone_period = steps;
    o->stop_rule = PMMC_STOP_FIRST_FAIL;
  }
  switch (level) {
    case PMMC_LEVEL_HEATMAP:
      o->guard_energy = 1e-6;  // ParallelMonteCarloHeatmap.cu:465; LooperSolver.hpp:1465
      if (preset == PMMC_PRESET_CPU) {
        o->step = 0.5f * step_size;  // LooperSolver.hpp:1426
        o->small_energy = 1e-6;      // :1550
      } else if (preset == PMMC_PRESET_GPU) {
        o->step = 0.75f * step_size;  // ParallelMonteCarloHeatmap.cu:478
        o->step_decay = 0.95f;        // :286
        o->block = 512;               // :247
        o->milestone_period = 512;
        o->cool_per_block = 1;        // :285
        o->stop_rule = PMMC_STOP_MISSES_ACCUM;
        o->min_successes = INT_MAX;   // misses counted on E > thr*ms alone (:343-345)
        o->small_energy = 1e-4;       // :349
      } else {
        return PMMC_ERR_INVALID;
      }
      return PMMC_OK;
    case PMMC_LEVEL_ARCS:
      o->step = step_size;
      o->small_energy = 1e-5;  // LooperSolver.hpp:4778; .cu:641
      o->ratio_stop = 0.9999;  // :4779; .cu:642
      if (preset == PMMC_PRESET_GPU) {
        o->block = 256;  // ParallelMonteCarloArcs.cu:520
        o->milestone_period = 256;
        o->stop_rule = PMMC_STOP_MISSES_RESET;
        o->positive_guard = 1;  // .cu:599
        o->max_moves = steps;   // .cu:522
      } else if (preset != PMMC_PRESET_CPU) {
        return PMMC_ERR_INVALID;
      }
      return PMMC_OK;
    case PMMC_LEVEL_SMOOTH:
      o->step = step_size;
      o->small_energy = 1e-6;  // LooperSolver.hpp:5030; .cu:325
      o->strict_accept = 1;    // LooperSolver.hpp:4974; .cu:297
      if (preset == PMMC_PRESET_GPU) {
        o->block = 256;  // ParallelMonteCarloSmooth.cu:269
        o->milestone_period = steps;  // .cu:320
        o->stop_rule = PMMC_STOP_MISSES_ACCUM;
        o->max_moves = steps;  // .cu:267
      } else if (preset != PMMC_PRESET_CPU) {
        return PMMC_ERR_INVALID;
      }
      return PMMC_OK;
    default:
      return PMMC_ERR_INVALID;
  }
}

namespace pmmc {

pmmc_status check_schedule(pmmc_ctx *ctx, const pmmc_schedule &s) {
  if (s.block <= 0 || s.milestone_period <= 0) return fail(ctx, PMMC_ERR_INVALID, "schedule: block and milestone_period must be > 0");
  if (s.max_moves < 0 || s.max_moves > 0xFFFFFFFFll - s.block) return fail(ctx, PMMC_ERR_INVALID, "schedule: max_moves out of range");
  if (!(s.dt > 0.0) || !std::isfinite(s.T0)) return fail(ctx, PMMC_ERR_INVALID, "schedule: bad temperature");
  if (!std::isfinite(s.step) || s.step < 0.f) return fail(ctx, PMMC_ERR_INVALID, "schedule: bad step");
  if (s.stop_rule < 0 || s.stop_rule > 2) return fail(ctx, PMMC_ERR_INVALID, "schedule: bad stop rule");
  return PMMC_OK;
}

int default_chains(const pmmc_ctx *ctx, const HostProblem &h) {
  // Whole rounds of the level kernel's resident wavefronts: the ensembles
  // (64 chains each) that one CU holds at once, times the rounds needed for
  // at least sixteen per CU -- enough wavefronts to hide the HBM latency of
  // the partner loads, and no partial last round (a partial round runs its
  // few waves alone, which measured 10-20% slower) -- capped so the chains'
  // positions take at most a third of the device's HBM (96 GB on MI355X; a
  // 20,000-bead smooth problem needs 78 GB for one full round).
  const int cus = ctx && ctx->cu_count > 0 ? ctx->cu_count : 256;
  const int res = resident_ensembles_per_cu(h);
  const int rounds = (16 + res - 1) / res;
  const size_t per_ens = static_cast<size_t>(h.n > 0 ? h.n : 1) * 3 * kLanes * sizeof(float);
  size_t ens = static_cast<size_t>(res) * rounds * cus;
  const size_t hbm = ctx && ctx->hbm_bytes > 0 ? ctx->hbm_bytes : (size_t(96) << 30) * 3;
  const size_t cap = (hbm / 3) / per_ens;
  if (ens > cap) ens = cap > 0 ? cap : 1;
  return static_cast<int>(ens) * kLanes;
}

// heatmap_chromosome_boundaries lookup, LooperSolver.hpp:6034-6050
static void block_of(const int32_t *b, int nb, int n, int p, int &st, int &en) {
  st = 0;
  en = n - 1;
  if (nb <= 0 || !b) return;
  if (p < 0 || p > b[nb - 1]) return;
  for (int i = 0; i + 1 < nb; ++i)
    if (b[i] <= p && p < b[i + 1]) {
      st = b[i];
      en = b[i + 1] - 1;
      return;
    }
}

static pmmc_status prep_heatmap(pmmc_ctx *ctx, const pmmc_heatmap_problem *p, HostProblem &h) {
  const int n = p->n;
  if (n < 1 || !p->pos || !p->dist) return fail(ctx, PMMC_ERR_INVALID, "heatmap: n < 1 or null arrays");
  if (p->diagonal < 0 || p->n_boundaries < 0 || (p->n_boundaries > 0 && !p->boundaries))
    return fail(ctx, PMMC_ERR_INVALID, "heatmap: bad diagonal or boundaries");
  const size_t nn = static_cast<size_t>(n) * n;
  bool nan = false;
  for (size_t i = 0; i < nn; ++i) nan |= std::isnan(p->dist[i]);
  if (nan) return fail(ctx, PMMC_ERR_INVALID, "heatmap: NaN in dist");
  std::vector<int> st(n), en(n);
  for (int m = 0; m < n; ++m) block_of(p->boundaries, p->n_boundaries, n, m, st[m], en[m]);
  const float thr = p->skip_below;
  const int diag = p->diagonal;
  // rows padded with zeros to 32 / 64 entries for small heatmaps, so the
  // register-resident kernel reads whole rows with wide scalar loads
  h.ldw = n <= 32 ? 32 : (n <= 64 ? 64 : n);
  const int ldw = h.ldw;
  const size_t nl = static_cast<size_t>(n) * ldw;
  h.W.assign(nl, 0.f);
  h.Hc.assign(nl, 0.f);
  std::vector<float> Hc2(nl, 0.f);
  h.W2.assign(nl, 0.f);
  // W2 / Hc2 row q: the term of bead q in local(j), valid when q lies in j's
  // chromosome block, off the diagonal band and at or above the threshold
  // (row q of dist, read in order; selects instead of branches)
  for (int q = 0; q < n; ++q) {
    const float *dr = p->dist + static_cast<size_t>(q) * n;
    float *w2 = h.W2.data() + static_cast<size_t>(q) * ldw;
    float *c2 = Hc2.data() + static_cast<size_t>(q) * ldw;
    for (int j = 0; j < n; ++j) {
      const float hv = dr[j];
      const int dq = q > j ? q - j : j - q;
      const bool ok = q >= st[j] && q <= en[j] && q != j && dq >= diag && hv >= thr;
      w2[j] = ok ? 1.0f / hv : 0.f;
      c2[j] = ok ? hv : 0.f;
    }
  }
  // W / Hc row q: the term of bead j in local(q) -- the same value as W2 /
  // Hc2 at (j, q), so W and Hc are their transposes (32 x 32 tiles)
  for (int q0 = 0; q0 < n; q0 += 32)
    for (int j0 = 0; j0 < n; j0 += 32)
      for (int q = q0; q < std::min(n, q0 + 32); ++q)
        for (int j = j0; j < std::min(n, j0 + 32); ++j) {
          h.W[static_cast<size_t>(q) * ldw + j] = h.W2[static_cast<size_t>(j) * ldw + q];
          h.Hc[static_cast<size_t>(q) * ldw + j] = Hc2[static_cast<size_t>(j) * ldw + q];
        }
  // The kernels' symmetric form (ΔE = 2 * the row-p terms, re-sums over i < j
  // doubled) is exact only when the term of partner j in local(p) equals the
  // term of p in local(j) for every pair.  That depends on the targets AND
  // on the validity rule (diagonal band, skip threshold, chromosome blocks:
  // boundaries that do not partition [0, n) give some beads the whole range
  // and others one block), so it is decided on the built tables.
  h.symmetric = std::memcmp(h.W.data(), h.W2.data(), sizeof(float) * nl) == 0 &&
                std::memcmp(h.Hc.data(), Hc2.data(), sizeof(float) * nl) == 0;
  if (h.symmetric) std::vector<float>().swap(h.W2);
  h.pos0.assign(p->pos, p->pos + 3 * static_cast<size_t>(n));
  h.fixed.assign(n, 0);
  if (p->fixed) h.fixed.assign(p->fixed, p->fixed + n);
  return PMMC_OK;
}

static pmmc_status prep_arcs(pmmc_ctx *ctx, const pmmc_arcs_problem *p, HostProblem &h) {
  const int n = p->n;
  if (n < 1 || !p->pos || !p->exp_dist) return fail(ctx, PMMC_ERR_INVALID, "arcs: n < 1 or null arrays");
  if (p->n_loops < 0 || (p->n_loops > 0 && (!p->loop_pairs || !p->loop_params)))
    return fail(ctx, PMMC_ERR_INVALID, "arcs: bad loop arrays");
  const float *ed = p->exp_dist;
  std::vector<std::vector<ArcEntry>> rows(n);
  for (int i = 0; i < n; ++i) {
    const float *row = ed + static_cast<size_t>(i) * n;
    for (int j0 = 0; j0 < n; j0 += 64) {
      const int j1 = std::min(n, j0 + 64);
      // nearly every entry is negative (no arc): a run of 64 with no
      // non-negative or NaN entry is passed over after one vector test
      bool any = false;
      for (int j = j0; j < j1; ++j) any |= !(row[j] < 0.f);
      if (!any) continue;
      for (int j = j0; j < j1; ++j) {
        const float e = row[j];
        if (std::isnan(e)) return fail(ctx, PMMC_ERR_INVALID, "arcs: NaN in exp_dist");
        if (j == i) continue;
        if (e >= 1e-6f) {
          const float e2 = e * e;
          rows[i].push_back(ArcEntry{j, e, p->k_stretch / e2, p->k_squeeze / e2});
        } else if (e >= 0.f && j > i) {
          h.zpairs.push_back(i);
          h.zpairs.push_back(j);
        }
      }
    }
  }
  for (int k = 0; k < p->n_loops; ++k) {
    const int a = p->loop_pairs[2 * k], b = p->loop_pairs[2 * k + 1];
    if (a < 0 || a >= n || b < 0 || b >= n) return fail(ctx, PMMC_ERR_INVALID, "arcs: loop index out of range");
    if (a == b) continue;  // distance stays 0: contributes a constant, no delta
    const float kk = p->loop_params[2 * k], r0 = p->loop_params[2 * k + 1];
    rows[a].push_back(ArcEntry{static_cast<int32_t>(b | INT32_MIN), r0, kk, kk});
    rows[b].push_back(ArcEntry{static_cast<int32_t>(a | INT32_MIN), r0, kk, kk});
  }
  h.off.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) h.off[i + 1] = h.off[i] + static_cast<int32_t>(rows[i].size());
  h.ent.reserve(h.off[n]);
  for (int i = 0; i < n; ++i) h.ent.insert(h.ent.end(), rows[i].begin(), rows[i].end());
  for (const ArcEntry &e : h.ent) {
    h.ej.push_back(e.j);
    h.ec.push_back(e.c);
    h.eka.push_back(e.ka);
    h.ekb.push_back(e.kb);
  }
  for (int i = 0; i < n; ++i)
    for (const ArcEntry &e : rows[i])
      if ((e.j & INT32_MAX) > i) h.pairs.push_back(PairEntry{i, e});
  h.ks = p->k_stretch;
  h.kq = p->k_squeeze;
  h.pos0.assign(p->pos, p->pos + 3 * static_cast<size_t>(n));
  h.fixed.assign(n, 0);
  if (p->fixed) h.fixed.assign(p->fixed, p->fixed + n);
  return PMMC_OK;
}

static pmmc_status prep_smooth(pmmc_ctx *ctx, const pmmc_smooth_problem *p, HostProblem &h) {
  const int n = p->n;
  if (n < 1 || !p->pos || !p->dist_to_next) return fail(ctx, PMMC_ERR_INVALID, "smooth: n < 1 or null arrays");
  h.dtn.assign(p->dist_to_next, p->dist_to_next + n);
  h.ks = p->k_stretch;
  h.kq = p->k_squeeze;
  h.ka = p->k_angular;
  h.wd = p->w_dist;
  h.wa = p->w_angle;
  h.pos0.assign(p->pos, p->pos + 3 * static_cast<size_t>(n));
  h.fixed.assign(n, 0);
  if (p->fixed) h.fixed.assign(p->fixed, p->fixed + n);
  return PMMC_OK;
}

static pmmc_status prepare_tables(pmmc_ctx *ctx, int level, const void *problem,
                                  const pmmc_schedule &s, const pmmc_run &run, HostProblem &h) {
  if (!problem) return fail(ctx, PMMC_ERR_INVALID, "null problem");
  pmmc_status ps = check_schedule(ctx, s);
  if (ps != PMMC_OK) return ps;
  h.level = level;
  switch (level) {
    case PMMC_LEVEL_HEATMAP: ps = prep_heatmap(ctx, static_cast<const pmmc_heatmap_problem *>(problem), h); break;
    case PMMC_LEVEL_ARCS: ps = prep_arcs(ctx, static_cast<const pmmc_arcs_problem *>(problem), h); break;
    case PMMC_LEVEL_SMOOTH: ps = prep_smooth(ctx, static_cast<const pmmc_smooth_problem *>(problem), h); break;
    default: return fail(ctx, PMMC_ERR_INVALID, "unknown level");
  }
  if (ps != PMMC_OK) return ps;
  h.n = static_cast<int>(h.pos0.size() / 3);
  for (int i = 0; i < h.n; ++i)
    if (!h.fixed[i]) h.free_list.push_back(i);
  return set_run(ctx, s, run, h);
}

// The host tables of a large problem are several n x n arrays: a failed host
// allocation becomes PMMC_ERR_NOMEM, never an exception through the C ABI
// (or out of a preparing thread, which would end the process).
pmmc_status prepare_problem(pmmc_ctx *ctx, int level, const void *problem, const pmmc_schedule &s,
                            const pmmc_run &run, HostProblem &h) {
  try {
    return prepare_tables(ctx, level, problem, s, run, h);
  } catch (const std::bad_alloc &) {
    h = HostProblem{};
    return fail(ctx, PMMC_ERR_NOMEM, "host memory exhausted while building the problem's tables");
  } catch (const std::length_error &) {
    h = HostProblem{};
    return fail(ctx, PMMC_ERR_INVALID, "problem too large for its host tables");
  }
}

pmmc_status set_run(pmmc_ctx *ctx, const pmmc_schedule &s, const pmmc_run &run, HostProblem &h) {
  pmmc_status ps = check_schedule(ctx, s);
  if (ps != PMMC_OK) return ps;
  h.sched = s;
  h.seed = run.seed ? run.seed : ctx->default_seed;
  h.segment = run.segment;
  if (run.chain0 % kLanes != 0) return fail(ctx, PMMC_ERR_INVALID, "run.chain0 must be a multiple of 64");
  h.chain0 = run.chain0;
  h.n_chains = run.n_chains > 0 ? run.n_chains : default_chains(ctx, h);
  if (static_cast<uint64_t>(h.chain0) + h.n_chains > 0x7FFFFFFFull)
    return fail(ctx, PMMC_ERR_INVALID, "chain ids exceed 2^31");
  h.pick_shift = pick_shift_of(h.level, h.n);
  h.workers = 0;
  if (h.level == PMMC_LEVEL_HEATMAP && hreg_halves(h) > 0 && !stage_ok(h.n)) {
    // queued register kernels: at most one resident round of worker
    // workgroups; their slots take the remaining chains as chains stop
    const int n_ens = (h.n_chains + kLanes - 1) / kLanes;
    const int cus = ctx && ctx->cu_count > 0 ? ctx->cu_count : 256;
    h.workers = std::min(n_ens, std::max(1, resident_ensembles_per_cu(h) * cus));
  }
  return PMMC_OK;
}

}  // namespace pmmc
