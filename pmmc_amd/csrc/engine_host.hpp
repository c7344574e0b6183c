// engine_host.hpp -- host-side problem preparation for the annealer.
#pragma once

#include <stdint.h>

#include <tuple>
#include <vector>

#include "ctx.hpp"
#include "engine.hpp"

namespace pmmc {

// A validated problem with its derived device tables, still in host memory.
struct HostProblem {
  int level = 0, n = 0;
  std::vector<int32_t> free_list;
  std::vector<uint8_t> fixed;
  std::vector<float> pos0;
  // heatmap tables (see DevProblem)
  std::vector<float> W, W2, Hc;
  bool symmetric = true;
  int ldw = 0;  // row stride of W, W2, Hc
  // arcs CSR
  std::vector<int32_t> off;
  std::vector<ArcEntry> ent;
  // the same entries as separate arrays (partner | loop<<31, c, ka, kb)
  std::vector<int32_t> ej;
  std::vector<float> ec, eka, ekb;
  std::vector<PairEntry> pairs;
  std::vector<int32_t> zpairs;
  float ks = 0.f, kq = 0.f;
  // smooth
  std::vector<float> dtn;
  float ka = 0.f, wd = 1.f, wa = 1.f;
  pmmc_schedule sched{};
  uint64_t seed = 0;
  uint32_t segment = 0, chain0 = 0;
  int n_chains = 0;
  // bead-pick groups (pick_shift_of) and, for the queued register kernels,
  // the worker workgroups (set by prepare_problem)
  int pick_shift = 6;
  int workers = 0;
};

// Bead-pick group of chain c is c >> pick_shift_of(level, n): the 64 chains
// of an ensemble share picks (6), except heatmaps of 33..64 beads, whose
// register-resident kernels draw them per chain (0).
int pick_shift_of(int level, int n);

pmmc_status prepare_problem(pmmc_ctx *ctx, int level, const void *problem, const pmmc_schedule &s,
                            const pmmc_run &run, HostProblem &out);
// The schedule and run fields of a prepared problem (the last step of
// prepare_problem): a copy of a prepared problem given another schedule and
// run equals prepare_problem of the same problem struct with them.
pmmc_status set_run(pmmc_ctx *ctx, const pmmc_schedule &s, const pmmc_run &run, HostProblem &h);
// prepare_problem of every job (on up to 16 host threads).  A job naming the
// same level and problem struct as an earlier job of the list copies that
// job's prepared tables and sets its own schedule and run, instead of building
// the tables again (an n x n scan per arc or heatmap problem).  On failure
// *failed is the first failing job and ctx's error message is that job's.
pmmc_status prepare_jobs(pmmc_ctx *ctx, const pmmc_job *jobs, int n_jobs, std::vector<HostProblem> &hp,
                         int *failed);
pmmc_status check_schedule(pmmc_ctx *ctx, const pmmc_schedule &s);
int default_chains(const pmmc_ctx *ctx, const HostProblem &h);
// ensembles (one wavefront each) of h's annealing kernel one CU holds at once
// (the runtime's occupancy figure for that kernel; 16 if the query fails)
int resident_ensembles_per_cu(const HostProblem &h);

// contiguous range of ensembles launched together with one kernel:
// (level, LDS staging and its n, register-resident heatmap halves, asymmetry)
struct LaunchGroup {
  int level;
  bool stage;
  int n;      // staged groups: the problems' n (sets the LDS size); else 0
  int hreg;   // heatmap kept in registers: 2 (33..64 beads, two 32-bead halves); 0 = no
  bool asym;  // register heatmap group holds a non-symmetric heatmap
  int ens0, n_ens;
  // arc level: chains per wavefront of the launch, 64 (one lane per chain) or
  // 32 (a lane pair per chain, arcs_pair.hip); the same results either way
  int cpw = 64;
};
LaunchGroup group_of(const HostProblem &h);
std::tuple<int, int, int, int, int> group_key(const LaunchGroup &g);

// everything a batch needs on the device
struct BatchDev {
  DevProblem *probs = nullptr;
  DevEns *ens = nullptr;
  DevResult *results = nullptr;
  int *counter = nullptr;
  std::vector<DevProblem> hprobs;
  std::vector<DevEns> hens;
  std::vector<LaunchGroup> groups;
  int n_ens_total = 0;
  std::vector<char> stage;      // staging of the static tables for a caller-given span
                                // (otherwise the context's pinned buffer)
  char *outs = nullptr;         // every problem's result positions
  size_t outs_bytes = 0;
  std::vector<size_t> out_off;  // problem i's positions at outs + out_off[i]
};

// Device memory a batch is carved from: a caller-given span (base non-null;
// nothing is allocated) or, when base is null, the context's arena.
struct DevSpan {
  char *base = nullptr;
  size_t cap = 0;
};
// device bytes a batch of these problems needs
size_t batch_bytes(const std::vector<HostProblem> &hp);
// Launch shape of a run: blocks per step launch (bps <= 0 without tune: one
// launch runs every chain to its end), steps per captured graph, hipGraph
// replay on/off, arc chains per wavefront (0 = not set); tune: fields left 0
// are chosen by timing the run's first steps (tuned = 1 afterwards).
struct BatchShape {
  int bps = 0, spg = 0, cpw = 0;
  bool graph = false, tune = false;
  int tuned = 0;
};
// Run prepared problems to completion with launch shape `sh` (in/out).
pmmc_status run_problems(pmmc_ctx *ctx, std::vector<HostProblem> &hp, std::vector<float *> &outs,
                         std::vector<pmmc_anneal_stats *> &stats, BatchShape &sh,
                         DevSpan span = DevSpan{});

bool stage_ok(int n);
size_t problem_bytes(const HostProblem &h);
// Upload one problem into carved device memory; its static tables are staged
// in `stage`, which must stay alive until the stream has synchronised.
void fill_problem(Carver &cv, hipStream_t s, const HostProblem &h, int index, DevResult *result,
                  DevProblem &P, std::vector<DevEns> &ens, hipError_t &err, std::vector<char> &stage);
hipError_t launch_groups(hipStream_t s, const DevProblem *probs, const DevEns *ens,
                         const std::vector<LaunchGroup> &groups, int max_blocks);
// max_n: largest bead count among the problems (sizes the replication grid)
hipError_t launch_init(hipStream_t s, DevProblem *probs, int n_probs, const DevEns *ens, int n_ens,
                       int max_n);
// the same for problems and ensembles addressed through references (server)
hipError_t launch_init_refs(hipStream_t s, const JobRef *jobs, int n_jobs, const EnsRef *ens,
                            int n_ens, int max_n);
hipError_t launch_finalize_refs(hipStream_t s, const JobRef *jobs, int n_jobs);
// register-resident heatmap kernel (heatmap_reg.hip) for heatmaps of <= 64 beads
int hreg_halves(const HostProblem &h);
int hreg_resident_per_cu(bool asym);
hipError_t launch_heatmap_reg(hipStream_t s, const DevProblem *probs, const DevEns *ens, int n_ens,
                              int max_blocks, int halves, bool asym);
hipError_t launch_finalize(hipStream_t s, DevProblem *probs, int n_probs, const DevEns *ens);
// lane-pair arc kernel (arcs_pair.hip): one 128-lane workgroup per ensemble
int arcs_pair_resident_per_cu();
hipError_t launch_arcs_pair(hipStream_t s, const DevProblem *probs, const DevEns *ens, int n_ens,
                            int max_blocks);
hipError_t launch_count_per_problem(hipStream_t s, const DevEns *ens, int n_ens, int *counts,
                                    int n_counts);

}  // namespace pmmc
