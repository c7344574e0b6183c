// server.hip -- continuous-batching annealing server.
//
// Jobs (any level, any size) are queued by pmmc_server_submit.  Each
// pmmc_server_step admits queued jobs while the resident set fits the HBM
// budget and the slot count, runs ONE launch per launch group over every
// resident chain (all jobs' chains overlap on the device), and retires jobs
// whose chains have all stopped.  A chain's trajectory depends only on its
// (seed, segment, chain id) Philox stream and its own state, which is saved
// and restored bit-exactly between launches, so a job's result is the same
// whether it runs alone or among others, and whatever number of moves each
// launch runs.
//
// Device memory: the HBM budget is allocated once, at creation, and carved by
// a host-side first-fit free list (256-byte granules), so admission and
// retirement never call hipMalloc / hipFree (hipFree would wait for the
// device).  Retirement finalizes every finished job of a step, copies their
// results into one pinned staging buffer and synchronizes once.
//
// Pacing: resident jobs are grouped by kernel (level, register/LDS variant)
// and by size class (floor(log2 n)).  With blocks_per_step = 0 each group's
// launch runs as many blocks of moves as fit a target (launch_ms, default
// kTargetMs) measured
// on the device (HIP events around the group's launch, the budget rescaled
// after every step), so small jobs finish in a few steps while large ones
// advance one block at a time.  The groups' launches of a step are spread
// over `streams` streams (default kStreams, all at the default priority):
// groups touch disjoint jobs, and a step's small launches would otherwise run
// one after another on an idle device.  The first is the library's stream,
// the next three the context's side streams (created with the context, so
// that each holds a hardware queue of its own even after a hipGraph launch
// has taken one); a server asked for more creates the rest itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <set>
#include <tuple>
#include <vector>

#include "ctx.hpp"
#include "engine_host.hpp"

namespace pmmc {

// Default device time per group launch when pacing adaptively.  On the
// served mix (Poisson 200 requests/s of heatmap, arc and smooth jobs), 8 s
// runs: p50 / p99 latency 206 / 515 ms at 4 ms, 177 / 424 at 8, 145 / 397 at
// 12, 143 / 428 at 16, 160 / 545 at 24, 199 / 378 at 48; 1 and 2 ms: 513 /
// 1,529 and 272 / 952.  Longer launches mean fewer steps (each pays the
// host's admission, synchronisation and retirement), until small jobs wait
// behind large launches.
constexpr float kTargetMs = 12.0f;
constexpr int kStreams = 4;         // default number of streams a step's group launches use
constexpr int kMaxStreams = 16;
constexpr int kMaxBlocks = 1 << 16;
constexpr size_t kGranule = 256;

struct SrvJob {
  uint64_t id = 0;
  HostProblem hp;
  float *pos_out = nullptr;
  pmmc_anneal_stats *stats = nullptr;
  size_t bytes = 0;
  size_t off = 0;               // offset of the job's block in the pool
  char *mem = nullptr;          // pool + off while resident
  DevProblem *dprob = nullptr;  // job-local problem (index 0)
  DevEns *dens = nullptr;       // job-local ensembles (problem = 0)
  DevResult *dres = nullptr;
  DevProblem hprob{};
  std::vector<DevEns> hens;
  std::vector<char> stage;      // the job's static tables as uploaded at admission
};

// launch group of the server: kernel group key + size class
using SrvKey = std::tuple<int, int, int, int, int, int>;

struct SrvGroup {
  LaunchGroup g;
  SrvKey key;
};

struct Pacing {
  int blocks = 4;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

// First-fit free list over [0, cap) in kGranule units.
class FreeList {
 public:
  void reset(size_t cap) {
    free_.clear();
    if (cap >= kGranule) free_[0] = cap / kGranule * kGranule;
  }
  bool alloc(size_t bytes, size_t *off) {
    const size_t need = (bytes + kGranule - 1) / kGranule * kGranule;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < need) continue;
      *off = it->first;
      const size_t rest = it->second - need;
      const size_t at = it->first + need;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    size_t size = (bytes + kGranule - 1) / kGranule * kGranule;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + size == next->first) {  // merge with the next block
      size += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {  // merge into the previous block
        prev->second += size;
        return;
      }
    }
    free_[off] = size;
  }

 private:
  std::map<size_t, size_t> free_;
};

}  // namespace pmmc

struct pmmc_server {
  pmmc_ctx *ctx = nullptr;
  pmmc_server_opts opts{};
  uint64_t next_id = 1;
  std::deque<pmmc::SrvJob *> queue;
  std::vector<pmmc::SrvJob *> resident;
  std::set<uint64_t> done;
  size_t resident_bytes = 0;
  bool dirty = true;
  char *pool = nullptr;
  pmmc::FreeList free_list;
  pmmc::DevProblem *d_probs = nullptr;
  size_t probs_cap = 0;
  pmmc::DevEns *d_ens = nullptr;
  size_t ens_cap = 0;
  int *d_counts = nullptr;
  size_t counts_cap = 0;
  char *h_stage = nullptr;  // pinned staging of retired results
  size_t stage_cap = 0;
  std::vector<pmmc::SrvGroup> groups;
  std::vector<pmmc::DevProblem> h_probs;  // host images of d_probs / d_ens (kept
  std::vector<pmmc::DevEns> h_ens;        // alive while their copies are queued)
  pmmc::JobRef *d_jobrefs = nullptr;      // jobs initialised / finalised by one launch
  size_t jobrefs_cap = 0;
  pmmc::EnsRef *d_ensrefs = nullptr;
  size_t ensrefs_cap = 0;
  std::vector<pmmc::JobRef> h_jobrefs;
  std::vector<pmmc::EnsRef> h_ensrefs;
  std::map<pmmc::SrvKey, pmmc::Pacing> pacing;
  int n_ens = 0;
  // the streams besides ctx->stream (the context's side streams, then any
  // the server created itself, which it owns) and the events that order
  // them with it: `ready` after the step's uploads, joined[k] after stream
  // k's launches
  std::vector<hipStream_t> side;
  std::vector<hipStream_t> own;
  hipEvent_t ready = nullptr;
  std::vector<hipEvent_t> joined;
};

using namespace pmmc;

static size_t job_bytes(const HostProblem &h) {
  const int ne = (h.n_chains + kLanes - 1) / kLanes;
  return problem_bytes(h) + 3 * 256 + sizeof(DevProblem) + sizeof(DevEns) * ne + sizeof(DevResult);
}

static int size_class(int n) {
  int c = 0;
  while ((2 << c) <= n) ++c;
  return c;
}

static SrvKey server_key(const HostProblem &h) {
  const auto k = group_key(group_of(h));
  return std::make_tuple(std::get<0>(k), std::get<1>(k), std::get<2>(k), std::get<3>(k),
                         std::get<4>(k), size_class(h.n));
}

template <class T>
static pmmc_status grow(pmmc_ctx *ctx, T *&p, size_t &cap, size_t need) {
  if (need <= cap) return PMMC_OK;
  if (p) {
    PMMC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(p);
    p = nullptr;
  }
  size_t c = std::max<size_t>(need, 2 * cap);
  if (hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * c) != hipSuccess)
    return fail(ctx, PMMC_ERR_NOMEM, "server: allocation failed");
  cap = c;
  return PMMC_OK;
}

// Place the job in the pool and upload it; false when the pool has no
// contiguous block for it yet (the job stays queued).
static pmmc_status admit(pmmc_server *s, SrvJob *j, bool *placed) {
  pmmc_ctx *ctx = s->ctx;
  *placed = s->free_list.alloc(j->bytes, &j->off);
  if (!*placed) return PMMC_OK;
  j->mem = s->pool + j->off;
  Carver cv{j->mem, 0, j->bytes};
  const int ne = (j->hp.n_chains + kLanes - 1) / kLanes;
  j->dprob = cv.take<DevProblem>(1);
  j->dens = cv.take<DevEns>(ne);
  j->dres = cv.take<DevResult>(1);
  hipError_t err = hipSuccess;
  j->hens.clear();
  fill_problem(cv, ctx->stream, j->hp, 0, j->dres, j->hprob, j->hens, err, j->stage);
  if (err == hipSuccess)
    err = hipMemcpyAsync(j->dprob, &j->hprob, sizeof(DevProblem), hipMemcpyHostToDevice, ctx->stream);
  if (err == hipSuccess)
    err = hipMemcpyAsync(j->dens, j->hens.data(), sizeof(DevEns) * ne, hipMemcpyHostToDevice, ctx->stream);
  if (err != hipSuccess) {
    s->free_list.release(j->off, j->bytes);
    j->mem = nullptr;
    *placed = false;
    return fail(ctx, PMMC_ERR_HIP, std::string("server admit: ") + hipGetErrorString(err));
  }
  return PMMC_OK;
}

// Rebuild the combined problem / ensemble arrays and the launch groups of the
// resident set.
static pmmc_status rebuild(pmmc_server *s) {
  pmmc_ctx *ctx = s->ctx;
  std::vector<DevProblem> &probs = s->h_probs;
  probs.clear();
  std::vector<std::pair<SrvKey, int>> keys;
  for (size_t k = 0; k < s->resident.size(); ++k) {
    keys.emplace_back(server_key(s->resident[k]->hp), static_cast<int>(k));
    probs.push_back(s->resident[k]->hprob);
  }
  std::stable_sort(keys.begin(), keys.end(),
                   [](const std::pair<SrvKey, int> &a, const std::pair<SrvKey, int> &b) {
                     return a.first < b.first;
                   });
  std::vector<DevEns> &ens = s->h_ens;
  ens.clear();
  s->groups.clear();
  for (auto &t : keys) {
    const int k = t.second;
    if (s->groups.empty() || s->groups.back().key != t.first) {
      SrvGroup sg;
      sg.g = group_of(s->resident[k]->hp);
      sg.g.ens0 = static_cast<int>(ens.size());
      sg.g.n_ens = 0;
      sg.key = t.first;
      s->groups.push_back(sg);
    }
    for (DevEns e : s->resident[k]->hens) {
      e.problem = k;
      ens.push_back(e);
      s->groups.back().g.n_ens++;
    }
  }
  s->n_ens = static_cast<int>(ens.size());
  pmmc_status ps = grow(ctx, s->d_probs, s->probs_cap, std::max<size_t>(1, probs.size()));
  if (ps == PMMC_OK) ps = grow(ctx, s->d_ens, s->ens_cap, std::max<size_t>(1, ens.size()));
  if (ps == PMMC_OK) ps = grow(ctx, s->d_counts, s->counts_cap, std::max<size_t>(1, probs.size()));
  if (ps != PMMC_OK) return ps;
  if (!probs.empty())
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(s->d_probs, probs.data(), sizeof(DevProblem) * probs.size(),
                                     hipMemcpyHostToDevice, ctx->stream));
  if (!ens.empty())
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(s->d_ens, ens.data(), sizeof(DevEns) * ens.size(),
                                     hipMemcpyHostToDevice, ctx->stream));
  s->dirty = false;
  return PMMC_OK;
}

// Finalize the finished jobs, copy their results through the pinned staging
// buffer with one synchronization, write them to the callers' buffers and
// return the jobs' pool blocks.
static pmmc_status retire(pmmc_server *s, const std::vector<SrvJob *> &fin) {
  if (fin.empty()) return PMMC_OK;
  pmmc_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  std::vector<size_t> at(fin.size());
  size_t need = 0;
  for (size_t i = 0; i < fin.size(); ++i) {
    at[i] = need;
    need += (sizeof(DevResult) + sizeof(float) * fin[i]->hp.pos0.size() + 15) & ~size_t(15);
  }
  if (need > s->stage_cap) {
    PMMC_HIP_TRY(ctx, hipStreamSynchronize(st));
    if (s->h_stage) (void)hipHostFree(s->h_stage);
    s->h_stage = nullptr;
    const size_t cap = std::max(need, 2 * s->stage_cap);
    if (hipHostMalloc(reinterpret_cast<void **>(&s->h_stage), cap) != hipSuccess) {
      s->stage_cap = 0;
      return fail(ctx, PMMC_ERR_NOMEM, "server: pinned staging allocation failed");
    }
    s->stage_cap = cap;
  }
  s->h_jobrefs.clear();
  for (SrvJob *j : fin) s->h_jobrefs.push_back(JobRef{j->dprob, j->dens});
  pmmc_status ps = grow(ctx, s->d_jobrefs, s->jobrefs_cap, s->h_jobrefs.size());
  if (ps != PMMC_OK) return ps;
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(s->d_jobrefs, s->h_jobrefs.data(), sizeof(JobRef) * s->h_jobrefs.size(),
                                   hipMemcpyHostToDevice, st));
  PMMC_HIP_TRY(ctx, launch_finalize_refs(st, s->d_jobrefs, static_cast<int>(fin.size())));
  for (size_t i = 0; i < fin.size(); ++i) {
    SrvJob *j = fin[i];
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(s->h_stage + at[i], j->dres, sizeof(DevResult),
                                     hipMemcpyDeviceToHost, st));
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(s->h_stage + at[i] + sizeof(DevResult), j->hprob.out_pos,
                                     sizeof(float) * j->hp.pos0.size(), hipMemcpyDeviceToHost, st));
  }
  PMMC_HIP_TRY(ctx, hipStreamSynchronize(st));
  for (size_t i = 0; i < fin.size(); ++i) {
    SrvJob *j = fin[i];
    DevResult R;
    std::memcpy(&R, s->h_stage + at[i], sizeof(R));
    const float *tmp = reinterpret_cast<const float *>(s->h_stage + at[i] + sizeof(DevResult));
    if (j->pos_out) {
      if (j->hp.level == PMMC_LEVEL_SMOOTH) {
        // only non-fixed beads are written back (ParallelMonteCarloSmooth.cu:485-491)
        for (int b = 0; b < j->hp.n; ++b)
          if (!j->hp.fixed[b])
            for (int k = 0; k < 3; ++k) j->pos_out[3 * b + k] = tmp[3 * b + k];
      } else {
        std::memcpy(j->pos_out, tmp, sizeof(float) * j->hp.pos0.size());
      }
    }
    if (j->stats) {
      pmmc_anneal_stats &S = *j->stats;
      S.energy = R.energy;
      S.energy_resummed = R.resummed;
      S.best_energy = R.best;
      S.moves = R.moves;
      S.accepted = R.accepted;
      S.best_chain_moves = R.best_moves;
      S.best_chain_accepted = R.best_accepted;
      S.best_chain = R.best_chain;
      S.n_chains = R.n_chains;
      S.milestones = R.milestones;
      S.stop_reason = R.stop_reason;
      S.kernel_ms = 0.0;
    }
    s->free_list.release(j->off, j->bytes);
    j->mem = nullptr;
    s->resident_bytes -= j->bytes;
    s->done.insert(j->id);
  }
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_server_create(pmmc_ctx *ctx, const pmmc_server_opts *opts,
                                          pmmc_server **out) {
  if (!ctx || !opts || !out) return fail(ctx, PMMC_ERR_INVALID, "server: null argument");
  if (opts->max_jobs <= 0 || opts->hbm_budget == 0 || opts->blocks_per_step < 0)
    return fail(ctx, PMMC_ERR_INVALID,
                "server: max_jobs and hbm_budget must be > 0, blocks_per_step >= 0");
  if (!(opts->launch_ms >= 0.f) || opts->launch_ms > 1000.f || opts->streams < 0 ||
      opts->streams > kMaxStreams)
    return fail(ctx, PMMC_ERR_INVALID, "server: launch_ms must be in [0, 1000], streams in [0, 16]");
  pmmc_server *s = new pmmc_server();
  s->ctx = ctx;
  s->opts = *opts;
  if (s->opts.launch_ms == 0.f) s->opts.launch_ms = kTargetMs;
  if (s->opts.streams == 0) s->opts.streams = kStreams;
  if (hipMalloc(reinterpret_cast<void **>(&s->pool), s->opts.hbm_budget) != hipSuccess) {
    delete s;
    return fail(ctx, PMMC_ERR_NOMEM, "server: the HBM budget could not be allocated");
  }
  s->free_list.reset(s->opts.hbm_budget);
  bool ok = hipEventCreateWithFlags(&s->ready, hipEventDisableTiming) == hipSuccess;
  for (int k = 1; ok && k < s->opts.streams; ++k) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    if (static_cast<size_t>(k - 1) < ctx->side.size()) {
      st = ctx->side[k - 1];
    } else {
      ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
      if (ok) s->own.push_back(st);
    }
    if (ok) s->side.push_back(st);
    ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
    if (ok) s->joined.push_back(ev);
  }
  if (!ok) {
    (void)hipGetLastError();
    pmmc_server_destroy(s);
    return fail(ctx, PMMC_ERR_HIP, "server: stream or event creation failed");
  }
  *out = s;
  return PMMC_OK;
}

extern "C" void pmmc_server_destroy(pmmc_server *s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->ctx->stream);
  for (SrvJob *j : s->resident) delete j;
  for (SrvJob *j : s->queue) delete j;
  for (auto &kv : s->pacing) {
    if (kv.second.e0) (void)hipEventDestroy(kv.second.e0);
    if (kv.second.e1) (void)hipEventDestroy(kv.second.e1);
  }
  if (s->pool) (void)hipFree(s->pool);
  if (s->d_probs) (void)hipFree(s->d_probs);
  if (s->d_ens) (void)hipFree(s->d_ens);
  if (s->d_counts) (void)hipFree(s->d_counts);
  if (s->d_jobrefs) (void)hipFree(s->d_jobrefs);
  if (s->d_ensrefs) (void)hipFree(s->d_ensrefs);
  if (s->h_stage) (void)hipHostFree(s->h_stage);
  for (hipStream_t st : s->side) (void)hipStreamSynchronize(st);
  for (hipStream_t st : s->own) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : s->joined) (void)hipEventDestroy(ev);
  if (s->ready) (void)hipEventDestroy(s->ready);
  delete s;
}

static pmmc_status submit_many(pmmc_server *s, const pmmc_job *jobs, int32_t n_jobs,
                               uint64_t *job_ids) {
  for (int i = 0; i < n_jobs; ++i)
    if (!jobs[i].pos_out) return fail(s->ctx, PMMC_ERR_INVALID, "server: bad job");
  // every job is prepared and sized before any is queued, so a failing call
  // queues nothing
  std::vector<HostProblem> hp;
  int bad = -1;
  pmmc_status ps = prepare_jobs(s->ctx, jobs, n_jobs, hp, &bad);
  if (ps != PMMC_OK) return ps;
  std::vector<size_t> bytes(n_jobs);
  for (int i = 0; i < n_jobs; ++i) {
    bytes[i] = job_bytes(hp[i]);
    if (bytes[i] > s->opts.hbm_budget) return fail(s->ctx, PMMC_ERR_NOMEM, "server: job larger than the HBM budget");
  }
  std::vector<std::unique_ptr<SrvJob>> made(n_jobs);
  for (int i = 0; i < n_jobs; ++i) {
    made[i].reset(new SrvJob());
    made[i]->hp = std::move(hp[i]);
    made[i]->bytes = bytes[i];
    made[i]->pos_out = jobs[i].pos_out;
    made[i]->stats = jobs[i].stats;
  }
  const size_t queued = s->queue.size();
  try {
    for (auto &j : made) s->queue.push_back(j.get());
  } catch (...) {
    s->queue.resize(queued);
    throw;
  }
  for (int i = 0; i < n_jobs; ++i) {
    made[i]->id = s->next_id++;
    if (job_ids) job_ids[i] = made[i]->id;
    made[i].release();
  }
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_server_submit_many(pmmc_server *s, const pmmc_job *jobs, int32_t n_jobs,
                                               uint64_t *job_ids) {
  if (!s || (!jobs && n_jobs > 0) || n_jobs < 0)
    return fail(s ? s->ctx : nullptr, PMMC_ERR_INVALID, "server: bad job list");
  return no_throw(s->ctx, [&] { return submit_many(s, jobs, n_jobs, job_ids); });
}

extern "C" pmmc_status pmmc_server_submit(pmmc_server *s, const pmmc_job *job, uint64_t *job_id) {
  if (!s || !job || !job->pos_out) return fail(s ? s->ctx : nullptr, PMMC_ERR_INVALID, "server: bad job");
  return pmmc_server_submit_many(s, job, 1, job_id);
}

// Initial totals and chain state of the jobs admitted in this step: one
// launch pair over all of them.
static pmmc_status init_jobs(pmmc_server *s, const std::vector<SrvJob *> &fresh) {
  pmmc_ctx *ctx = s->ctx;
  s->h_jobrefs.clear();
  s->h_ensrefs.clear();
  int max_n = 1;
  for (SrvJob *j : fresh) {
    s->h_jobrefs.push_back(JobRef{j->dprob, j->dens});
    for (size_t e = 0; e < j->hens.size(); ++e) s->h_ensrefs.push_back(EnsRef{j->dprob, j->dens + e});
    max_n = std::max(max_n, j->hp.n);
  }
  pmmc_status ps = grow(ctx, s->d_jobrefs, s->jobrefs_cap, s->h_jobrefs.size());
  if (ps == PMMC_OK) ps = grow(ctx, s->d_ensrefs, s->ensrefs_cap, s->h_ensrefs.size());
  if (ps != PMMC_OK) return ps;
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(s->d_jobrefs, s->h_jobrefs.data(), sizeof(JobRef) * s->h_jobrefs.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(s->d_ensrefs, s->h_ensrefs.data(), sizeof(EnsRef) * s->h_ensrefs.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
  PMMC_HIP_TRY(ctx, launch_init_refs(ctx->stream, s->d_jobrefs, static_cast<int>(s->h_jobrefs.size()),
                                     s->d_ensrefs, static_cast<int>(s->h_ensrefs.size()), max_n));
  return PMMC_OK;
}

// One launch per group, round-robin over the streams (each side stream first
// waits for the step's uploads on ctx->stream; ctx->stream then waits for
// every side stream); with adaptive pacing each launch is bracketed by the
// group's events on its stream.
static pmmc_status launch_step(pmmc_server *s) {
  pmmc_ctx *ctx = s->ctx;
  const bool adaptive = s->opts.blocks_per_step == 0;
  const int n_side = std::min<int>(static_cast<int>(s->side.size()),
                                   std::max<int>(0, static_cast<int>(s->groups.size()) - 1));
  if (n_side > 0) {
    PMMC_HIP_TRY(ctx, hipEventRecord(s->ready, ctx->stream));
    for (int k = 0; k < n_side; ++k) PMMC_HIP_TRY(ctx, hipStreamWaitEvent(s->side[k], s->ready, 0));
  }
  for (size_t i = 0; i < s->groups.size(); ++i) {
    const SrvGroup &sg = s->groups[i];
    const int slot = n_side > 0 ? static_cast<int>(i % static_cast<size_t>(n_side + 1)) : 0;
    hipStream_t st = slot == 0 ? ctx->stream : s->side[slot - 1];
    int blocks = s->opts.blocks_per_step;
    Pacing *pc = nullptr;
    if (adaptive) {
      pc = &s->pacing[sg.key];
      if (!pc->e0) {
        PMMC_HIP_TRY(ctx, hipEventCreate(&pc->e0));
        PMMC_HIP_TRY(ctx, hipEventCreate(&pc->e1));
      }
      blocks = pc->blocks;
      PMMC_HIP_TRY(ctx, hipEventRecord(pc->e0, st));
    }
    std::vector<LaunchGroup> one(1, sg.g);
    PMMC_HIP_TRY(ctx, launch_groups(st, s->d_probs, s->d_ens, one, blocks));
    if (pc) PMMC_HIP_TRY(ctx, hipEventRecord(pc->e1, st));
  }
  for (int k = 0; k < n_side; ++k) {
    PMMC_HIP_TRY(ctx, hipEventRecord(s->joined[k], s->side[k]));
    PMMC_HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, s->joined[k], 0));
  }
  return PMMC_OK;
}

// After the step's synchronization: rescale each group's block budget toward
// launch_ms of device time.
static void repace(pmmc_server *s) {
  if (s->opts.blocks_per_step != 0) return;
  for (const SrvGroup &sg : s->groups) {
    Pacing &pc = s->pacing[sg.key];
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pc.e0, pc.e1) != hipSuccess || !(ms > 0.f)) continue;
    const double scale = static_cast<double>(s->opts.launch_ms) / ms;
    double nb = pc.blocks * std::min(4.0, std::max(0.25, scale));
    pc.blocks = static_cast<int>(std::min<double>(kMaxBlocks, std::max(1.0, std::floor(nb))));
  }
}

extern "C" pmmc_status pmmc_server_step(pmmc_server *s, int32_t *n_done, int32_t *n_resident,
                                        int32_t *n_queued) {
  if (!s) return PMMC_ERR_INVALID;
  pmmc_ctx *ctx = s->ctx;
  int finished = 0;
  std::vector<SrvJob *> fresh;
  // admission: FIFO, while slots, budget and a contiguous pool block allow.
  // A job whose upload fails is dropped (its block returned); the jobs
  // admitted before it are still initialised, then the step reports the error.
  pmmc_status admit_err = PMMC_OK;
  while (!s->queue.empty() && static_cast<int>(s->resident.size()) < s->opts.max_jobs &&
         s->resident_bytes + s->queue.front()->bytes <= s->opts.hbm_budget) {
    SrvJob *j = s->queue.front();
    bool placed = false;
    admit_err = admit(s, j, &placed);
    if (admit_err != PMMC_OK) {
      s->queue.pop_front();
      delete j;
      break;
    }
    if (!placed) break;
    s->queue.pop_front();
    s->resident.push_back(j);
    fresh.push_back(j);
    s->resident_bytes += j->bytes;
    s->dirty = true;
  }
  if (!fresh.empty()) {
    pmmc_status ps = init_jobs(s, fresh);
    if (ps != PMMC_OK) return ps;
  }
  if (admit_err != PMMC_OK) return admit_err;
  if (!s->resident.empty()) {
    if (s->dirty) {
      pmmc_status ps = rebuild(s);
      if (ps != PMMC_OK) return ps;
    }
    pmmc_status ps = launch_step(s);
    if (ps != PMMC_OK) return ps;
    const int nr = static_cast<int>(s->resident.size());
    PMMC_HIP_TRY(ctx, launch_count_per_problem(ctx->stream, s->d_ens, s->n_ens, s->d_counts, nr));
    std::vector<int> counts(nr);
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(counts.data(), s->d_counts, sizeof(int) * nr, hipMemcpyDeviceToHost,
                                     ctx->stream));
    PMMC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    repace(s);
    std::vector<SrvJob *> keep, fin;
    for (int k = 0; k < nr; ++k) (counts[k] == 0 ? fin : keep).push_back(s->resident[k]);
    ps = retire(s, fin);
    if (ps != PMMC_OK) return ps;
    for (SrvJob *j : fin) delete j;
    finished = static_cast<int>(fin.size());
    if (!fin.empty()) s->dirty = true;
    s->resident.swap(keep);
  }
  if (n_done) *n_done = finished;
  if (n_resident) *n_resident = static_cast<int32_t>(s->resident.size());
  if (n_queued) *n_queued = static_cast<int32_t>(s->queue.size());
  return PMMC_OK;
}

extern "C" int pmmc_server_done(const pmmc_server *s, uint64_t job_id) {
  return s && s->done.count(job_id) ? 1 : 0;
}
