// arcs_pair.hip -- arc-level annealer with a chain split over a lane pair.
//
// The same Markov chains, random streams, schedule and results as the
// one-lane-per-chain arc path of k_anneal (engine.hip), bit for bit: the
// arc ΔE is defined as the sum of the even CSR entries' terms plus the sum
// of the odd entries' terms (kernels.hpp, arcs_term), and here lane 2i of a
// wavefront forms the even sum of chain i and lane 2i + 1 the odd sum; one
// DPP swap gives both lanes even + odd.  Everything else (random draws, the
// FP64 acceptance, the schedule) is computed identically in both lanes.
// A 128-lane workgroup runs one 64-chain ensemble (wave w: chains 32w ..
// 32w + 31), so a problem with few chains occupies twice the wavefronts of
// the one-lane form -- twice the moves in flight where the one-lane form
// leaves the GPU latency-bound (batches of many problems with a few hundred
// chains each).  Only the moved bead's own position is read by both lanes.
#include <hip/hip_runtime.h>

#include "ctx.hpp"
#include "engine_host.hpp"
#include "kernels.hpp"

namespace pmmc {

constexpr int kPairChunk = kArcsChunk / 2;  // a lane's entries of the first chunk

// Move m+1 fetched ahead, the lane's half: entries e0 + half + 2u of the row.
struct ArcsFetch2 {
  int p, e0, e1;
  int j[kPairChunk];
  float px, py, pz;
  float xj[kPairChunk], yj[kPairChunk], zj[kPairChunk];
};

// Move mw + k of its window (kernels.hpp, ArcsWindow: lane k gathered the
// move's pick and row indices); this lane takes entries half, half + 2, ...
__device__ __forceinline__ void arcs_fetch2(const PosView &X, const ArcsWindow &w, int k, int half,
                                            ArcsFetch2 &f) {
  f.p = __builtin_amdgcn_readlane(w.p, k);
  f.e0 = __builtin_amdgcn_readlane(w.e0, k);
  f.e1 = __builtin_amdgcn_readlane(w.e1, k);
  f.px = X.get(f.p, 0);
  f.py = X.get(f.p, 1);
  f.pz = X.get(f.p, 2);
  if (f.e1 > f.e0) {
#pragma unroll
    for (int u = 0; u < kPairChunk; ++u) {
      const int even = __builtin_amdgcn_readlane(w.j[2 * u], k);
      const int odd = __builtin_amdgcn_readlane(w.j[2 * u + 1], k);
      f.j[u] = half ? odd : even;
    }
#pragma unroll
    for (int u = 0; u < kPairChunk; ++u) {
      f.xj[u] = X.get(f.j[u], 0);
      f.yj[u] = X.get(f.j[u], 1);
      f.zj[u] = X.get(f.j[u], 2);
    }
  }
}

// the current move changed bead q (per chain: ok): patch the fetched copies
__device__ __forceinline__ void arcs_patch2(ArcsFetch2 &f, int half, int q, bool ok, float qx,
                                            float qy, float qz) {
  const bool self = ok && f.p == q;
  f.px = self ? qx : f.px;
  f.py = self ? qy : f.py;
  f.pz = self ? qz : f.pz;
  const int cnt = f.e1 - f.e0;
#pragma unroll
  for (int u = 0; u < kPairChunk; ++u) {
    const bool hit = ok && half + 2 * u < cnt && f.j[u] == q;
    f.xj[u] = hit ? qx : f.xj[u];
    f.yj[u] = hit ? qy : f.yj[u];
    f.zj[u] = hit ? qz : f.zj[u];
  }
}

// ΔE of a fetched move: this lane's parity sum, then even + odd in both lanes
__device__ __forceinline__ float arcs_delta2(const DevProblem &P, const PosView &X,
                                             const ArcsFetch2 &f, int half, float nx, float ny,
                                             float nz) {
  float acc = 0.f;
  if (f.e1 > f.e0) {
    // the lane's entries of the first chunk, loaded before any term (the
    // tables carry kArcsChunk entries of padding past the last row)
    float c[kPairChunk], ka[kPairChunk], kb[kPairChunk];
#pragma unroll
    for (int u = 0; u < kPairChunk; ++u) {
      const int e = f.e0 + half + 2 * u;
      c[u] = P.ec[e];
      ka[u] = P.eka[e];
      kb[u] = P.ekb[e];
    }
#pragma unroll
    for (int u = 0; u < kPairChunk; ++u) {
      if (f.e0 + half + 2 * u < f.e1)
        acc += arcs_term(c[u], ka[u], kb[u], f.px, f.py, f.pz, nx, ny, nz, f.xj[u], f.yj[u],
                         f.zj[u]);
    }
  }
  // entries past the first chunk (beads of more than kArcsChunk entries)
  for (int e = f.e0 + kArcsChunk + half; e < f.e1; e += 2) {
    const int j = P.ej[e] & 0x7fffffff;
    acc += arcs_term(P.ec[e], P.eka[e], P.ekb[e], f.px, f.py, f.pz, nx, ny, nz, X.get(j, 0),
                     X.get(j, 1), X.get(j, 2));
  }
  // the pair's other sum: DPP quad_perm [1,0,3,2] swaps adjacent lanes
  const float other =
      __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(acc), 0xB1, 0xF, 0xF, false));
  return acc + other;  // IEEE addition commutes: both lanes hold even + odd
}

__global__ __launch_bounds__(128) void k_anneal_arcs2(const DevProblem *__restrict__ probs,
                                                      const DevEns *__restrict__ ens,
                                                      int max_blocks) {
  const DevEns it = ens[blockIdx.x];
  const DevProblem &P = probs[it.problem];
  const int half = threadIdx.x & 1;
  const int cl = threadIdx.x >> 1;  // chain within the ensemble
  EnsState *__restrict__ st = it.st;
  int status = st->status[cl];
  if (__ballot(status == 0) == 0) return;  // no barrier in this kernel: a wave may leave

  const PosView X = global_view(it.pos, static_cast<unsigned>(cl));
  const pmmc_schedule S = P.sched;
  const uint64_t seed = P.seed;
  const uint32_t segment = P.segment;
  const float r = S.sphere_radius;
  const double offset = *P.offset;

  double E = st->E[cl], T = st->T[cl], ms = st->ms[cl], best = st->best[cl];
  float step = st->step[cl];
  uint32_t iters = st->iters[cl], acc = st->acc[cl];
  int msucc = st->msucc[cl], misses = st->misses[cl], miles = st->milestones[cl];
  bool active = (status == 0);
  const uint32_t chain = it.chain0 + static_cast<uint32_t>(cl);
  const uint32_t group = it.chain0 >> kPickShift;

  for (int b = 0; b < max_blocks; ++b) {
    if (active && static_cast<int64_t>(iters) >= S.max_moves) {
      status = PMMC_STOP_CAP;
      active = false;
    }
    const uint64_t am = __ballot(active);
    if (am == 0) break;
    const uint32_t m0 = readlane_u32(iters, __builtin_ctzll(am));
    // software-pipelined as the one-lane path: move m+1's row and positions
    // are fetched before move m is evaluated, then patched
    ArcsWindow w;
    arcs_window(P, seed, group, segment, m0, w);
    ArcsFetch2 f;
    arcs_fetch2(X, w, 0, half, f);
    for (int k = 0; k < S.block; ++k) {
      const uint32_t m = m0 + static_cast<uint32_t>(k);
      const ArcsFetch2 c = f;
      const bool more = k + 1 < S.block;
      if (more) {
        const int kn = k + 1;
        if ((kn & (kLanes - 1)) == 0) arcs_window(P, seed, group, segment, m + 1, w);
        arcs_fetch2(X, w, kn & (kLanes - 1), half, f);
      }
      const U4 rc = draw(seed, m, chain, segment, PMMC_LEVEL_ARCS, kSubChain);
      const float dx = disp(rc.x, step), dy = disp(rc.y, step);
      const float dz = S.use2d ? 0.f : disp(rc.z, step);
      const float nx = move_coord(c.px, dx), ny = move_coord(c.py, dy), nz = move_coord(c.pz, dz);
      float dE = arcs_delta2(P, X, c, half, nx, ny, nz);
      if (r > 0.f) dE += sphere_penalty(nx, ny, nz, r) - sphere_penalty(c.px, c.py, c.pz, r);
      const double En = E + static_cast<double>(dE);
      const bool ok = active && metropolis(S, E, En, T, u01(rc.w));
      if (ok) {
        if (half == 0) {
          X.set(c.p, 0, nx);
          X.set(c.p, 1, ny);
          X.set(c.p, 2, nz);
        }
        E = En;
        ++msucc;
        ++acc;
        best = E < best ? E : best;
      }
      if (more) arcs_patch2(f, half, c.p, ok, nx, ny, nz);
      if (active && !S.cool_per_block) T *= S.dt;
    }
    bool due = false;
    if (active) {
      iters += static_cast<uint32_t>(S.block);
      if (S.cool_per_block) T *= S.dt;
      step *= S.step_decay;
      due = (iters % static_cast<uint32_t>(S.milestone_period)) < static_cast<uint32_t>(S.block);
    }
    if (__ballot(due) != 0) {
      double Er = E;
      const bool resum_now = S.resum > 0 && ((miles + 1) % S.resum) == 0;
      // both lanes re-sum the whole pair list (the one-lane path's order)
      if (__ballot(due && resum_now) != 0) Er = resum_arcs(P, X, r, offset);
      if (due) {
        ++miles;
        E = Er;
        milestone(S, E, ms, msucc, misses, status, active);
      }
    }
  }
  if (status == -1 || half != 0) return;  // padding chain, or the pair's second lane
  st->E[cl] = E;
  st->T[cl] = T;
  st->ms[cl] = ms;
  st->best[cl] = best;
  st->step[cl] = step;
  st->iters[cl] = iters;
  st->acc[cl] = acc;
  st->msucc[cl] = msucc;
  st->misses[cl] = misses;
  st->milestones[cl] = miles;
  st->status[cl] = status;
}

int arcs_pair_resident_per_cu() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_anneal_arcs2, 2 * kLanes, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return blocks;
}

hipError_t launch_arcs_pair(hipStream_t s, const DevProblem *probs, const DevEns *ens, int n_ens,
                            int max_blocks) {
  if (n_ens <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_anneal_arcs2, dim3(n_ens), dim3(2 * kLanes), 0, s, probs, ens, max_blocks);
  return hipGetLastError();
}

}  // namespace pmmc
