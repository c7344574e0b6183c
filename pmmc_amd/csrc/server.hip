// server.hip -- continuous-batching annealing server.
//
// Jobs (any level, any size) are queued by pmmc_server_submit.  Each
// pmmc_server_step admits queued jobs while the resident set fits the HBM
// budget and the slot count, runs ONE step launch per launch group over every
// resident chain (all jobs' chains overlap on the device), and retires jobs
// whose chains have all stopped.  A chain's trajectory depends only on its
// (seed, segment, chain id) Philox stream and its own state, which is saved
// and restored bit-exactly between steps, so a job's result is the same
// whether it runs alone or among others.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <set>
#include <tuple>
#include <vector>

#include "ctx.hpp"
#include "engine_host.hpp"

namespace pmmc {

struct SrvJob {
  uint64_t id = 0;
  HostProblem hp;
  float *pos_out = nullptr;
  pmmc_anneal_stats *stats = nullptr;
  size_t bytes = 0;
  char *mem = nullptr;
  DevProblem *dprob = nullptr;  // job-local problem (index 0)
  DevEns *dens = nullptr;       // job-local ensembles (problem = 0)
  DevResult *dres = nullptr;
  DevProblem hprob{};
  std::vector<DevEns> hens;
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
  pmmc::DevProblem *d_probs = nullptr;
  size_t probs_cap = 0;
  pmmc::DevEns *d_ens = nullptr;
  size_t ens_cap = 0;
  int *d_counts = nullptr;
  size_t counts_cap = 0;
  std::vector<pmmc::LaunchGroup> groups;
  int n_ens = 0;
};

using namespace pmmc;

static size_t job_bytes(const HostProblem &h) {
  const int ne = (h.n_chains + kLanes - 1) / kLanes;
  return problem_bytes(h) + 3 * 256 + sizeof(DevProblem) + sizeof(DevEns) * ne + sizeof(DevResult);
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

static pmmc_status admit(pmmc_server *s, SrvJob *j) {
  pmmc_ctx *ctx = s->ctx;
  if (hipMalloc(reinterpret_cast<void **>(&j->mem), j->bytes) != hipSuccess)
    return fail(ctx, PMMC_ERR_NOMEM, "server: job allocation failed");
  Carver cv{j->mem, 0, j->bytes};
  const int ne = (j->hp.n_chains + kLanes - 1) / kLanes;
  j->dprob = cv.take<DevProblem>(1);
  j->dens = cv.take<DevEns>(ne);
  j->dres = cv.take<DevResult>(1);
  hipError_t err = hipSuccess;
  j->hens.clear();
  fill_problem(cv, ctx->stream, j->hp, 0, j->dres, j->hprob, j->hens, err);
  if (err == hipSuccess)
    err = hipMemcpyAsync(j->dprob, &j->hprob, sizeof(DevProblem), hipMemcpyHostToDevice, ctx->stream);
  if (err == hipSuccess)
    err = hipMemcpyAsync(j->dens, j->hens.data(), sizeof(DevEns) * ne, hipMemcpyHostToDevice, ctx->stream);
  if (err == hipSuccess) err = launch_init(ctx->stream, j->dprob, 1, j->dens, ne);
  if (err != hipSuccess) return fail(ctx, PMMC_ERR_HIP, std::string("server admit: ") + hipGetErrorString(err));
  return PMMC_OK;
}

// Rebuild the combined problem / ensemble arrays of the resident set.
static pmmc_status rebuild(pmmc_server *s) {
  pmmc_ctx *ctx = s->ctx;
  std::vector<DevProblem> probs;
  std::vector<std::pair<std::tuple<int, int, int, int, int>, int>> keys;  // (group key, job)
  for (size_t k = 0; k < s->resident.size(); ++k) {
    keys.emplace_back(group_key(group_of(s->resident[k]->hp)), static_cast<int>(k));
    probs.push_back(s->resident[k]->hprob);
  }
  std::stable_sort(keys.begin(), keys.end());
  std::vector<DevEns> ens;
  s->groups.clear();
  for (auto &t : keys) {
    const int k = t.second;
    const LaunchGroup g = group_of(s->resident[k]->hp);
    if (s->groups.empty() || group_key(s->groups.back()) != t.first) {
      s->groups.push_back(g);
      s->groups.back().ens0 = static_cast<int>(ens.size());
    }
    for (DevEns e : s->resident[k]->hens) {
      e.problem = k;
      ens.push_back(e);
      s->groups.back().n_ens++;
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

static pmmc_status retire(pmmc_server *s, SrvJob *j) {
  pmmc_ctx *ctx = s->ctx;
  hipStream_t st = ctx->stream;
  PMMC_HIP_TRY(ctx, launch_finalize(st, j->dprob, 1, j->dens));
  DevResult R;
  std::vector<float> tmp(j->hp.pos0.size());
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(&R, j->dres, sizeof(R), hipMemcpyDeviceToHost, st));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(tmp.data(), j->hprob.out_pos, sizeof(float) * tmp.size(),
                                   hipMemcpyDeviceToHost, st));
  PMMC_HIP_TRY(ctx, hipStreamSynchronize(st));
  if (j->pos_out) {
    if (j->hp.level == PMMC_LEVEL_SMOOTH) {
      for (int b = 0; b < j->hp.n; ++b)
        if (!j->hp.fixed[b])
          for (int k = 0; k < 3; ++k) j->pos_out[3 * b + k] = tmp[3 * b + k];
    } else {
      std::memcpy(j->pos_out, tmp.data(), sizeof(float) * tmp.size());
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
  (void)hipFree(j->mem);
  j->mem = nullptr;
  s->resident_bytes -= j->bytes;
  s->done.insert(j->id);
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_server_create(pmmc_ctx *ctx, const pmmc_server_opts *opts,
                                          pmmc_server **out) {
  if (!ctx || !opts || !out) return fail(ctx, PMMC_ERR_INVALID, "server: null argument");
  if (opts->max_jobs <= 0 || opts->hbm_budget == 0)
    return fail(ctx, PMMC_ERR_INVALID, "server: max_jobs and hbm_budget must be > 0");
  pmmc_server *s = new pmmc_server();
  s->ctx = ctx;
  s->opts = *opts;
  if (s->opts.blocks_per_step <= 0) s->opts.blocks_per_step = 4;
  *out = s;
  return PMMC_OK;
}

extern "C" void pmmc_server_destroy(pmmc_server *s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->ctx->stream);
  for (SrvJob *j : s->resident) {
    if (j->mem) (void)hipFree(j->mem);
    delete j;
  }
  for (SrvJob *j : s->queue) delete j;
  if (s->d_probs) (void)hipFree(s->d_probs);
  if (s->d_ens) (void)hipFree(s->d_ens);
  if (s->d_counts) (void)hipFree(s->d_counts);
  delete s;
}

extern "C" pmmc_status pmmc_server_submit(pmmc_server *s, const pmmc_job *job, uint64_t *job_id) {
  if (!s || !job || !job->pos_out) return fail(s ? s->ctx : nullptr, PMMC_ERR_INVALID, "server: bad job");
  SrvJob *j = new SrvJob();
  pmmc_status ps = prepare_problem(s->ctx, job->level, job->problem, job->schedule, job->run, j->hp);
  if (ps != PMMC_OK) {
    delete j;
    return ps;
  }
  j->bytes = job_bytes(j->hp);
  if (j->bytes > s->opts.hbm_budget) {
    delete j;
    return fail(s->ctx, PMMC_ERR_NOMEM, "server: job larger than the HBM budget");
  }
  j->id = s->next_id++;
  j->pos_out = job->pos_out;
  j->stats = job->stats;
  s->queue.push_back(j);
  if (job_id) *job_id = j->id;
  return PMMC_OK;
}

extern "C" pmmc_status pmmc_server_step(pmmc_server *s, int32_t *n_done, int32_t *n_resident,
                                        int32_t *n_queued) {
  if (!s) return PMMC_ERR_INVALID;
  pmmc_ctx *ctx = s->ctx;
  int finished = 0;
  // admission: FIFO, while slots and budget allow
  while (!s->queue.empty() && static_cast<int>(s->resident.size()) < s->opts.max_jobs &&
         s->resident_bytes + s->queue.front()->bytes <= s->opts.hbm_budget) {
    SrvJob *j = s->queue.front();
    s->queue.pop_front();
    pmmc_status ps = admit(s, j);
    if (ps != PMMC_OK) {
      delete j;
      return ps;
    }
    s->resident.push_back(j);
    s->resident_bytes += j->bytes;
    s->dirty = true;
  }
  if (!s->resident.empty()) {
    if (s->dirty) {
      pmmc_status ps = rebuild(s);
      if (ps != PMMC_OK) return ps;
    }
    PMMC_HIP_TRY(ctx, launch_groups(ctx->stream, s->d_probs, s->d_ens, s->groups, s->opts.blocks_per_step));
    const int nr = static_cast<int>(s->resident.size());
    PMMC_HIP_TRY(ctx, launch_count_per_problem(ctx->stream, s->d_ens, s->n_ens, s->d_counts, nr));
    std::vector<int> counts(nr);
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(counts.data(), s->d_counts, sizeof(int) * nr, hipMemcpyDeviceToHost,
                                     ctx->stream));
    PMMC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<SrvJob *> keep;
    for (int k = 0; k < nr; ++k) {
      SrvJob *j = s->resident[k];
      if (counts[k] == 0) {
        pmmc_status ps = retire(s, j);
        if (ps != PMMC_OK) return ps;
        delete j;
        ++finished;
        s->dirty = true;
      } else {
        keep.push_back(j);
      }
    }
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
