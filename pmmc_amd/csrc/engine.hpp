// engine.hpp -- device data layout of the lockstep-ensemble annealer.
//
// Layout in HBM:
//   * one ensemble = 64 chains of one problem = one wavefront; lane = chain.
//   * ensemble positions are SoA  pos[(3*bead + axis) * 64 + lane]  (FP32), so
//     every position access of a wavefront is one coalesced 256-byte row.
//   * ensemble state is SoA over the 64 lanes (EnsState, 3.75 KB).
//   * per-problem static data (heatmap weight rows, arc CSR, dist_to_next,
//     free-bead list) is read with wave-uniform addresses, i.e. scalar loads.
#pragma once

#include <stdint.h>

#include "../../include/pmmc_hip.h"

namespace pmmc {

constexpr int kLanes = 64;
constexpr int kPickShift = 6;      // chains c and c' share bead picks iff c>>6 == c'>>6
constexpr int kResumMoves = 8192;  // presets re-sum the running total at most this often

// Philox level field of each annealer.
constexpr uint32_t kLevelHeatmap = 1, kLevelArcs = 2, kLevelSmooth = 3, kLevelSweep = 4;

struct EnsState {
  double E[kLanes];     // running total (FP64)
  double T[kLanes];     // temperature (FP64)
  double ms[kLanes];    // energy at the previous milestone
  double best[kLanes];  // lowest running total
  float step[kLanes];
  uint32_t iters[kLanes];
  uint32_t acc[kLanes];
  int32_t msucc[kLanes];
  int32_t misses[kLanes];
  int32_t milestones[kLanes];
  int32_t status[kLanes];  // 0 running, >0 pmmc_stop_reason, -1 unused lane
};

// arc CSR entry: partner | (loop << 31), centre c, stretch and squeeze
// coefficients (arcs: k/c^2, loops: k)
struct ArcEntry {
  int32_t j;
  float c, ka, kb;
};

// one arc or loop spring counted once (i < j), for the milestone re-sum
struct PairEntry {
  int32_t i;
  ArcEntry e;
};

struct DevProblem {
  int32_t level, n, nfree, symmetric;
  int32_t ldw, pad_ldw;      // row stride of W, W2, Hc (n, or 32/64 for small heatmaps)
  const int32_t *free_list;  // [nfree]
  const float *pos0;         // [n*3] initial positions (AoS)
  // heatmap: W[p*n+j] = valid(j,p) ? 1/h[j][p] : 0 ; W2 likewise with h[p][j]
  // (only for non-symmetric heatmaps); Hc[p*n+j] = valid(j,p) ? h[j][p] : 0
  const float *W, *W2, *Hc;
  // arcs
  const int32_t *off;        // [n+1]
  const ArcEntry *ent;       // [nnz]
  const int32_t *ej;         // [nnz] the entries' fields as separate arrays:
  const float *ec, *eka, *ekb;  // partner | loop<<31, c, ka, kb
  const int32_t *zpairs;     // [2*nz] off-diagonal pairs with 0 <= e < 1e-6 (i<j)
  const PairEntry *pairs;    // [npairs] the CSR entries with i < j, row order
  int32_t nnz, nz, npairs, pad2;
  float ks, kq;
  // smooth
  const float *dtn;
  float ka, wd, wa, pad0;
  // derived at init (device written)
  double *E0;                // [1] initial running total
  double *offset;            // [1] arcs: frozen repulsion; smooth: W0 - U0
  pmmc_schedule sched;
  uint64_t seed;
  uint32_t segment, chain0;
  int32_t ens0, n_ens, n_chains, pad1;
  float *out_pos;            // [n*3]
  struct DevResult *result;
  // chain queue of the register-resident heatmap kernels (33..64 beads): the
  // first `workers` ensembles' workgroups are workers whose 64 slots each run
  // one chain at a time and take the next chain from *queue when it stops
  int32_t workers;           // 0: every ensemble's workgroup runs its own chains
  int32_t pick_shift;        // chains c, c' share bead picks iff c>>s == c'>>s
  int32_t *queue;            // [1] next chain (relative to chain0) to hand out
  int32_t *slot_chain;       // [workers*64] chain a slot runs, -1 none, -2 drained
};

struct DevEns {
  int32_t problem;
  uint32_t chain0;  // chain id of lane 0 (multiple of 64)
  int32_t nlanes;
  int32_t pad;
  float *pos;       // [n*3*64]
  EnsState *st;
};

// a problem and the ensemble array its ens0 indexes (server jobs keep both
// in their own pool block)
struct JobRef {
  DevProblem *prob;
  const DevEns *ens;
};

// one ensemble of a problem, addressed directly
struct EnsRef {
  const DevProblem *prob;
  const DevEns *ens;
};

struct DevResult {
  double energy, resummed, best;
  int64_t moves, accepted, best_moves, best_accepted;
  int32_t best_chain, milestones, stop_reason, n_chains;
};

}  // namespace pmmc
