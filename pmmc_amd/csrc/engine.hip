// engine.hip -- lockstep-ensemble simulated annealing for the heatmap, arc
// and smooth levels of pMMC, plus the plain and batched C entry points.
//
// Reference kernels this replaces:
//   MonteCarloHeatmapKernel  ParallelMonteCarloHeatmap.cu:217-363 (chain form:
//                            the CPU annealer LooperSolver.hpp:1424-1565 on the GPU)
//   MonteCarloArcsKernel     ParallelMonteCarloArcs.cu:407-675
//   MonteCarloSmoothKernelMultiWarp ParallelMonteCarloSmooth.cu:231-349
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "ctx.hpp"
#include "engine_host.hpp"
#include "kernels.hpp"

namespace pmmc {

// ------------------------------------------------------------------ //
// FP64 exact energies (initial totals and the returned energies)
// ------------------------------------------------------------------ //
__device__ __forceinline__ double block_sum_d(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// heatmap full score (calcScoreHeatmapActiveRegion(), LooperSolver.hpp:3583-3617)
// over AoS positions, FP64 terms.
__device__ double exact_heatmap(const DevProblem &P, const float *x, double *red) {
  const int n = P.n;
  double s = 0.0;
  for (int m = threadIdx.x; m < n; m += blockDim.x) {
    const int j0 = P.symmetric ? m + 1 : 0;
    for (int j = j0; j < n; ++j) {
      const float h = P.Hc[static_cast<size_t>(m) * P.ldw + j];
      if (h == 0.f) continue;
      const double d = dist3d(x[3 * m], x[3 * m + 1], x[3 * m + 2], x[3 * j], x[3 * j + 1],
                              x[3 * j + 2]);
      const double c = (d - h) / h;
      s += c * c;
    }
  }
  double tot = block_sum_d(s, red);
  return P.symmetric ? 2.0 * tot : tot;
}

__device__ double exact_sphere(const DevProblem &P, const float *x, float r, double *red) {
  double s = 0.0;
  if (r > 0.f)
    for (int m = threadIdx.x; m < P.n; m += blockDim.x)
      s += sphere_penalty_d(x[3 * m], x[3 * m + 1], x[3 * m + 2], r);
  return block_sum_d(s, red);
}

// arc + loop terms, FP64 (LooperSolver.hpp:3215-3246 without repulsion, :3339-3353)
__device__ double exact_arcs_terms(const DevProblem &P, const float *x, double *red) {
  double s = 0.0;
  for (int p = threadIdx.x; p < P.n; p += blockDim.x) {
    for (int e = P.off[p]; e < P.off[p + 1]; ++e) {
      const ArcEntry en = P.ent[e];
      const int j = en.j & 0x7fffffff;
      if (j <= p) continue;
      const double d = dist3d(x[3 * p], x[3 * p + 1], x[3 * p + 2], x[3 * j], x[3 * j + 1],
                              x[3 * j + 2]);
      if (en.j < 0) {  // loop constraint
        const double t = d - en.c;
        s += static_cast<double>(en.ka) * t * t;
      } else {
        const double diff = (d - en.c) / en.c;
        s += diff * diff * (diff >= 0.0 ? P.ks : P.kq);
      }
    }
  }
  return block_sum_d(s, red);
}

// repulsion sum over pairs with exp_dist < 0 (LooperSolver.hpp:3228-3231):
// all pairs minus the pairs that carry an arc or an explicit zero.
__device__ double exact_repulsion(const DevProblem &P, const float *x, double *red) {
  const int n = P.n;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    for (int j = i + 1; j < n; ++j) {
      const double dx = xi - x[3 * j], dy = yi - x[3 * j + 1], dz = zi - x[3 * j + 2];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      s += 1.0 / (d > 1e-6 ? d : 1e-6);
    }
    for (int e = P.off[i]; e < P.off[i + 1]; ++e) {
      const ArcEntry en = P.ent[e];
      if (en.j < 0) continue;  // loops do not replace repulsion
      const int j = en.j;
      if (j <= i) continue;
      const double d = dist3d(x[3 * i], x[3 * i + 1], x[3 * i + 2], x[3 * j], x[3 * j + 1],
                              x[3 * j + 2]);
      s -= 1.0 / (d > 1e-6 ? d : 1e-6);
    }
  }
  for (int k = threadIdx.x; k < P.nz; k += blockDim.x) {
    const int i = P.zpairs[2 * k], j = P.zpairs[2 * k + 1];
    const double d = dist3d(x[3 * i], x[3 * i + 1], x[3 * i + 2], x[3 * j], x[3 * j + 1],
                            x[3 * j + 2]);
    s -= 1.0 / (d > 1e-6 ? d : 1e-6);
  }
  return block_sum_d(s, red);
}

// smooth bonds and angles, FP64 (LooperSolver.hpp:3414-3451); out[0] = bonds,
// out[1] = angles.
__device__ void exact_smooth(const DevProblem &P, const float *x, double *red, double *out) {
  const int n = P.n;
  double sb = 0.0, sa = 0.0;
  for (int i = threadIdx.x; i + 1 < n; i += blockDim.x) {
    double d0 = P.dtn[i];
    if (d0 < 1e-6) d0 = 1e-6;
    const double d = dist3d(x[3 * i], x[3 * i + 1], x[3 * i + 2], x[3 * i + 3], x[3 * i + 4],
                            x[3 * i + 5]);
    const double diff = (d - d0) / d0;
    sb += diff * diff * (diff >= 0.0 ? P.ks : P.kq);
    if (i >= 1) {
      const double ux = (double)x[3 * i - 3] - x[3 * i], uy = (double)x[3 * i - 2] - x[3 * i + 1],
                   uz = (double)x[3 * i - 1] - x[3 * i + 2];
      const double wx = (double)x[3 * i] - x[3 * i + 3], wy = (double)x[3 * i + 1] - x[3 * i + 4],
                   wz = (double)x[3 * i + 2] - x[3 * i + 5];
      const double lu = sqrt(ux * ux + uy * uy + uz * uz), lw = sqrt(wx * wx + wy * wy + wz * wz);
      if (lu > 1e-6 && lw > 1e-6) {
        double cs = (ux * wx + uy * wy + uz * wz) / (lu * lw);
        cs = cs < -1.0 ? -1.0 : (cs > 1.0 ? 1.0 : cs);
        const double ang = (1.0 - cs) * 0.5;
        sa += ang * ang * ang * P.ka;
      }
    }
  }
  out[0] = block_sum_d(sb, red);
  out[1] = block_sum_d(sa, red);
}

// Initial running total and offset of one problem (one 256-thread workgroup).
__device__ void init_problem_body(DevProblem &P, double *red) {
  const float *x = P.pos0;
  const float r = P.sched.sphere_radius;
  double E0 = 0.0, off = 0.0;
  if (P.level == PMMC_LEVEL_HEATMAP) {
    E0 = exact_heatmap(P, x, red) + exact_sphere(P, x, r, red);
  } else if (P.level == PMMC_LEVEL_ARCS) {
    off = exact_repulsion(P, x, red);
    E0 = off + exact_arcs_terms(P, x, red) + exact_sphere(P, x, r, red);
  } else {
    double bs[2];
    exact_smooth(P, x, red, bs);
    const double sph = exact_sphere(P, x, r, red);
    const double W0 = bs[0] * P.wd + bs[1] * P.wa + sph;
    const double U0 = bs[0] + bs[1] + sph;
    E0 = W0;
    off = W0 - U0;
  }
  if (threadIdx.x == 0) {
    *P.E0 = E0;
    *P.offset = off;
  }
  // chain queue: slot s starts on chain s, the queue hands out the rest
  if (P.workers > 0) {
    const int slots = P.workers * kLanes;
    for (int i = threadIdx.x; i < slots; i += blockDim.x) P.slot_chain[i] = i < P.n_chains ? i : -2;
    if (threadIdx.x == 0) *P.queue = slots;
  }
}

__global__ __launch_bounds__(256) void k_init_problem(DevProblem *probs) {
  __shared__ double red[256];
  init_problem_body(probs[blockIdx.x], red);
}

__global__ __launch_bounds__(256) void k_init_problem_refs(const JobRef *__restrict__ refs) {
  __shared__ double red[256];
  init_problem_body(*refs[blockIdx.x].prob, red);
}

// Ensemble initialisation: workgroup x of the ensemble replicates its share of
// pos0 into the SoA lanes (256 elements per pass, consecutive threads write
// consecutive addresses); workgroup 0 also sets the 64 lanes' state.
__device__ void init_ens_body(const DevProblem &P, const DevEns &it) {
  const size_t total = static_cast<size_t>(3 * P.n) * kLanes;
  // element i of the ensemble's positions: SoA rows hold pos0[i / 64];
  // lane-contiguous coordinates (smooth) hold bead i / 192, axis i % 3
  const bool xyz = P.level == PMMC_LEVEL_SMOOTH;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    it.pos[i] = xyz ? P.pos0[3 * (i / (3 * kLanes)) + i % 3] : P.pos0[i / kLanes];
  if (blockIdx.x != 0 || threadIdx.x >= kLanes) return;
  const int lane = threadIdx.x;
  EnsState *st = it.st;
  const double E0 = *P.E0;
  int status = 0;
  if (lane >= it.nlanes) status = -1;
  else if (!(E0 == E0) || E0 < P.sched.guard_energy) status = PMMC_STOP_GUARD;
  else if (P.nfree == 0) status = PMMC_STOP_NOFREE;
  st->E[lane] = E0;
  st->T[lane] = P.sched.T0;
  st->ms[lane] = E0;
  st->best[lane] = E0;
  st->step[lane] = P.sched.step;
  st->iters[lane] = 0;
  st->acc[lane] = 0;
  st->msucc[lane] = 0;
  st->misses[lane] = 0;
  st->milestones[lane] = 0;
  st->status[lane] = status;
}

__global__ __launch_bounds__(256) void k_init_ens(const DevProblem *__restrict__ probs,
                                                  const DevEns *__restrict__ ens, int n_ens) {
  for (int e = blockIdx.y; e < n_ens; e += gridDim.y) {
    const DevEns it = ens[e];
    init_ens_body(probs[it.problem], it);
  }
}

__global__ __launch_bounds__(256) void k_init_ens_refs(const EnsRef *__restrict__ refs, int n) {
  for (int e = blockIdx.y; e < n; e += gridDim.y) {
    const DevEns it = *refs[e].ens;
    init_ens_body(*refs[e].prob, it);
  }
}

// ------------------------------------------------------------------ //
// the annealing step kernel
// ------------------------------------------------------------------ //
template <int LEVEL, bool STAGE>
__global__ __launch_bounds__(64) void k_anneal(const DevProblem *__restrict__ probs,
                                               const DevEns *__restrict__ ens, int max_blocks) {
  extern __shared__ float lds[];
  const DevEns it = ens[blockIdx.x];
  const DevProblem &P = probs[it.problem];
  const int lane = threadIdx.x;
  EnsState *__restrict__ st = it.st;
  int status = st->status[lane];
  if (__ballot(status == 0) == 0) return;

  const int n = P.n;
  typename std::conditional<STAGE, PosViewT<true>, GlobalViewT<LEVEL>>::type X;
  if constexpr (STAGE) {
    for (int i = 0; i < 3 * n; ++i)
      lds[i * kLanes + lane] = it.pos[pos_index(kLaneXyz<LEVEL>, i / 3, i % 3, lane)];
    X.p = lds + lane;
  } else {
    X = level_view<LEVEL>(it.pos, lane);
  }

  const pmmc_schedule S = P.sched;
  const uint64_t seed = P.seed;
  const uint32_t segment = P.segment;
  const int nfree = P.nfree;
  const float r = S.sphere_radius;
  const double offset = (LEVEL == PMMC_LEVEL_HEATMAP) ? 0.0 : *P.offset;

  double E = st->E[lane], T = st->T[lane], ms = st->ms[lane], best = st->best[lane];
  float step = st->step[lane];
  uint32_t iters = st->iters[lane], acc = st->acc[lane];
  int msucc = st->msucc[lane], misses = st->misses[lane], miles = st->milestones[lane];
  bool active = (status == 0);
  const uint32_t chain = it.chain0 + static_cast<uint32_t>(lane);
  const uint32_t group = it.chain0 >> kPickShift;

  for (int b = 0; b < max_blocks; ++b) {
    if (active && static_cast<int64_t>(iters) >= S.max_moves) {
      status = PMMC_STOP_CAP;
      active = false;
    }
    const uint64_t am = __ballot(active);
    if (am == 0) break;
    const uint32_t m0 = readlane_u32(iters, __builtin_ctzll(am));
    if constexpr (LEVEL == PMMC_LEVEL_ARCS && !STAGE) {
      // Software-pipelined arc moves: the bead of move m+1 comes from the
      // pick stream, independent of the chains' state, so its CSR row and
      // the positions it reads are fetched before move m is evaluated (their
      // HBM latency overlaps move m's arithmetic); after move m's decision
      // the fetched copies of the bead move m changed are patched.  The
      // picks and row indices come from per-64-move windows (ArcsWindow).
      // Same draws, values and arithmetic as the plain loop below.
      constexpr int D = kArcsDepth;
      const int blk = S.block;
      ArcsFetch f[D];
      int kf = 0;  // the next move to fetch, relative to m0
      ArcsWindow w;
      arcs_window(P, seed, group, segment, m0, w);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (d < blk) {
          if (kf > 0 && (kf & (kLanes - 1)) == 0)
            arcs_window(P, seed, group, segment, m0 + static_cast<uint32_t>(kf), w);
          arcs_fetch_win(X, w, kf & (kLanes - 1), f[d]);
          ++kf;
        }
      }
      for (int k = 0; k < blk; ++k) {
        const uint32_t m = m0 + static_cast<uint32_t>(k);
        const ArcsFetch c = f[0];
#pragma unroll
        for (int d = 0; d + 1 < D; ++d) f[d] = f[d + 1];
        if (k + D < blk) {
          if ((kf & (kLanes - 1)) == 0)
            arcs_window(P, seed, group, segment, m0 + static_cast<uint32_t>(kf), w);
          arcs_fetch_win(X, w, kf & (kLanes - 1), f[D - 1]);
          ++kf;
        }
        const U4 rc = draw(seed, m, chain, segment, LEVEL, kSubChain);
        const float dx = disp(rc.x, step), dy = disp(rc.y, step);
        const float dz = S.use2d ? 0.f : disp(rc.z, step);
        const float nx = move_coord(c.px, dx), ny = move_coord(c.py, dy), nz = move_coord(c.pz, dz);
        float dE = arcs_delta_fetched(P, X, c, nx, ny, nz);
        if (r > 0.f) dE += sphere_penalty(nx, ny, nz, r) - sphere_penalty(c.px, c.py, c.pz, r);
        const double En = E + static_cast<double>(dE);
        const bool ok = active && metropolis(S, E, En, T, u01(rc.w));
        if (ok) {
          X.set(c.p, 0, nx);
          X.set(c.p, 1, ny);
          X.set(c.p, 2, nz);
          E = En;
          ++msucc;
          ++acc;
          best = E < best ? E : best;
        }
#pragma unroll
        for (int d = 0; d < D; ++d)
          if (k + 1 + d < blk) arcs_patch(f[d], c.p, ok, nx, ny, nz);
        if (active && !S.cool_per_block) T *= S.dt;
      }
    } else if constexpr (LEVEL == PMMC_LEVEL_SMOOTH && !STAGE) {
      // Software-pipelined smooth moves, as the arc moves above,
      // kSmoothDepth moves deep: the windows of moves m+1..m+kSmoothDepth
      // are in flight while move m is evaluated; after move m's decision
      // every fetched window is patched.
      // The picks and rest lengths come from per-64-move windows
      // (SmoothWindow); move m0 + i is lane i & 63 of window i >> 6.
      constexpr int D = kSmoothDepth;
      const int blk = S.block;
      SmoothFetch f[D];
      int kf = 0;  // the next move to fetch, relative to m0
      SmoothWindow w;
      smooth_window(P, seed, group, segment, m0, w);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (d < blk) {
          if (kf > 0 && (kf & (kLanes - 1)) == 0)
            smooth_window(P, seed, group, segment, m0 + static_cast<uint32_t>(kf), w);
          smooth_fetch_win(P, X, w, kf & (kLanes - 1), f[d]);
          ++kf;
        }
      }
      for (int k = 0; k < blk; ++k) {
        const uint32_t m = m0 + static_cast<uint32_t>(k);
        const SmoothFetch c = f[0];
#pragma unroll
        for (int d = 0; d + 1 < D; ++d) f[d] = f[d + 1];
        if (k + D < blk) {
          if ((kf & (kLanes - 1)) == 0)
            smooth_window(P, seed, group, segment, m0 + static_cast<uint32_t>(kf), w);
          smooth_fetch_win(P, X, w, kf & (kLanes - 1), f[D - 1]);
          ++kf;
        }
        const float px = c.qx[2], py = c.qy[2], pz = c.qz[2];
        const U4 rc = draw(seed, m, chain, segment, LEVEL, kSubChain);
        const float dx = disp(rc.x, step), dy = disp(rc.y, step);
        const float dz = S.use2d ? 0.f : disp(rc.z, step);
        const float nx = move_coord(px, dx), ny = move_coord(py, dy), nz = move_coord(pz, dz);
        float dE = smooth_delta(P, c.p, c.qx, c.qy, c.qz, px, py, pz, nx, ny, nz, c.d0m, c.d0p);
        if (r > 0.f) dE += sphere_penalty(nx, ny, nz, r) - sphere_penalty(px, py, pz, r);
        const double En = E + static_cast<double>(dE);
        const bool ok = active && metropolis(S, E, En, T, u01(rc.w));
        if (ok) {
          X.set(c.p, 0, nx);
          X.set(c.p, 1, ny);
          X.set(c.p, 2, nz);
          E = En;
          ++msucc;
          ++acc;
          best = E < best ? E : best;
        }
#pragma unroll
        for (int d = 0; d < D; ++d)
          if (k + 1 + d < blk) smooth_patch(f[d], c.p, ok, nx, ny, nz);
        if (active && !S.cool_per_block) T *= S.dt;
      }
    } else {
    U4 rb = {0, 0, 0, 0};
    for (int k = 0; k < S.block; ++k) {
      const uint32_t m = m0 + static_cast<uint32_t>(k);
      if (k == 0 || (m & 3u) == 0) rb = draw(seed, m >> 2, group, segment, LEVEL, kSubBead);
      const uint32_t sel = m & 3u;
      const uint32_t word = sel == 0 ? rb.x : (sel == 1 ? rb.y : (sel == 2 ? rb.z : rb.w));
      const int p = cst(P.free_list)[pick(word, static_cast<uint32_t>(nfree))];
      const float px = X.get(p, 0), py = X.get(p, 1), pz = X.get(p, 2);
      const U4 rc = draw(seed, m, chain, segment, LEVEL, kSubChain);
      const float dx = disp(rc.x, step), dy = disp(rc.y, step);
      const float dz = S.use2d ? 0.f : disp(rc.z, step);
      const float nx = move_coord(px, dx), ny = move_coord(py, dy), nz = move_coord(pz, dz);
      float dE;
      if constexpr (LEVEL == PMMC_LEVEL_HEATMAP) {
        dE = heatmap_delta(P, X, p, px, py, pz, nx, ny, nz);
      } else if constexpr (LEVEL == PMMC_LEVEL_ARCS) {
        dE = arcs_delta(P, X, p, px, py, pz, nx, ny, nz);
      } else {
        float qx[5], qy[5], qz[5];
#pragma unroll
        for (int o = 0; o < 5; ++o) {
          const int idx = p - 2 + o;
          const bool in = (o != 2) && idx >= 0 && idx < n;
          qx[o] = in ? X.get(idx, 0) : 0.f;
          qy[o] = in ? X.get(idx, 1) : 0.f;
          qz[o] = in ? X.get(idx, 2) : 0.f;
        }
        const float d0m = p > 0 ? fmaxf(cst(P.dtn)[p - 1], 1e-6f) : 1.f;
        const float d0p = p < n - 1 ? fmaxf(cst(P.dtn)[p], 1e-6f) : 1.f;
        dE = smooth_delta(P, p, qx, qy, qz, px, py, pz, nx, ny, nz, d0m, d0p);
      }
      if (r > 0.f) dE += sphere_penalty(nx, ny, nz, r) - sphere_penalty(px, py, pz, r);
      const double En = E + static_cast<double>(dE);
      const bool ok = active && metropolis(S, E, En, T, u01(rc.w));
      if (ok) {
        X.set(p, 0, nx);
        X.set(p, 1, ny);
        X.set(p, 2, nz);
        E = En;
        ++msucc;
        ++acc;
        best = E < best ? E : best;
      }
      if (active && !S.cool_per_block) T *= S.dt;
    }
    }
    bool due = false;
    if (active) {
      iters += static_cast<uint32_t>(S.block);
      if (S.cool_per_block) T *= S.dt;
      step *= S.step_decay;
      due = (iters % static_cast<uint32_t>(S.milestone_period)) < static_cast<uint32_t>(S.block);
    }
    // The active chains of an ensemble share `iters`, so they reach a
    // milestone together.  The re-summation runs with the whole wavefront
    // (it broadcasts table entries across lanes), stopped chains included;
    // only the chains at the milestone use its result.
    if (__ballot(due) != 0) {
      double Er = E;
      // due lanes share `iters`, hence `miles`: the re-sum decision is uniform
      const bool resum_now = S.resum > 0 && ((miles + 1) % S.resum) == 0;
      if (__ballot(due && resum_now) != 0) {
        if constexpr (LEVEL == PMMC_LEVEL_HEATMAP) Er = resum_heatmap(P, X, r);
        else if constexpr (LEVEL == PMMC_LEVEL_ARCS) Er = resum_arcs(P, X, r, offset);
        else Er = resum_smooth(P, X, r, offset);
      }
      if (due) {
        ++miles;
        E = Er;
        milestone(S, E, ms, msucc, misses, status, active);
      }
    }
  }
  if (status != -1) {
    st->E[lane] = E;
    st->T[lane] = T;
    st->ms[lane] = ms;
    st->best[lane] = best;
    st->step[lane] = step;
    st->iters[lane] = iters;
    st->acc[lane] = acc;
    st->msucc[lane] = msucc;
    st->misses[lane] = misses;
    st->milestones[lane] = miles;
    st->status[lane] = status;
  }
  if constexpr (STAGE) {
    for (int i = 0; i < 3 * n; ++i)
      it.pos[pos_index(kLaneXyz<LEVEL>, i / 3, i % 3, lane)] = X.p[i * kLanes];
  }
}

// Count lanes still running (for step loops).
__global__ __launch_bounds__(64) void k_count_active(const DevEns *__restrict__ ens, int n_ens,
                                                     int *count) {
  int c = 0;
  for (int e = blockIdx.x; e < n_ens; e += gridDim.x) c += ens[e].st->status[threadIdx.x] == 0;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if (threadIdx.x == 0 && c) atomicAdd(count, c);
}

__global__ __launch_bounds__(64) void k_zero_int(int *p) {
  if (threadIdx.x == 0) *p = 0;
}

// Moves done so far, summed over every chain (for the launch-shape tuner).
__global__ __launch_bounds__(64) void k_sum_iters(const DevEns *__restrict__ ens, int n_ens,
                                                  unsigned long long *sum) {
  unsigned long long c = 0;
  for (int e = blockIdx.x; e < n_ens; e += gridDim.x)
    if (threadIdx.x < ens[e].nlanes) c += ens[e].st->iters[threadIdx.x];
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if (threadIdx.x == 0 && c) atomicAdd(sum, c);
}

__global__ __launch_bounds__(64) void k_zero_u64(unsigned long long *p) {
  if (threadIdx.x == 0) *p = 0;
}

__global__ __launch_bounds__(64) void k_zero_n(int *p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// running lanes per problem index (counts zeroed first)
__global__ __launch_bounds__(64) void k_count_per_problem(const DevEns *__restrict__ ens, int n_ens,
                                                          int *counts) {
  for (int e = blockIdx.x; e < n_ens; e += gridDim.x) {
    int c = ens[e].st->status[threadIdx.x] == 0;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if (threadIdx.x == 0 && c) atomicAdd(&counts[ens[e].problem], c);
  }
}

hipError_t launch_count_per_problem(hipStream_t s, const DevEns *ens, int n_ens, int *counts,
                                    int n_counts) {
  hipLaunchKernelGGL(k_zero_n, dim3(std::max(1, std::min(64, (n_counts + 63) / 64))), dim3(64), 0, s,
                     counts, n_counts);
  if (n_ens > 0)
    hipLaunchKernelGGL(k_count_per_problem, dim3(std::min(n_ens, 1024)), dim3(64), 0, s, ens, n_ens,
                       counts);
  return hipGetLastError();
}

// Shared memory of the finalize workgroup.
struct FinalizeLds {
  double red[256];
  double bE[256];
  int bC[256];
  long long sm[256], sa[256];
};

// One workgroup (256 threads) per problem: best chain, AoS gather, FP64
// return value, move counts.  ens: the array P.ens0 indexes.
__device__ void finalize_body(DevProblem &P, const DevEns *__restrict__ ens, FinalizeLds &L) {
  double *red = L.red, *bE = L.bE;
  int *bC = L.bC;
  long long *sm = L.sm, *sa = L.sa;
  const int t = threadIdx.x;
  double myE = INFINITY;
  int myC = INT_MAX;
  long long mv = 0, ac = 0;
  const int total_lanes = P.n_ens * kLanes;
  for (int q = t; q < total_lanes; q += blockDim.x) {
    const DevEns &it = ens[P.ens0 + q / kLanes];
    const int lane = q % kLanes;
    if (lane >= it.nlanes) continue;
    const EnsState *st = it.st;
    const int c = static_cast<int>(it.chain0) + lane;
    mv += st->iters[lane];
    ac += st->acc[lane];
    const double e = st->E[lane];
    if (myC == INT_MAX || e < myE || (e == myE && c < myC)) {
      myE = e;
      myC = c;
    }
  }
  bE[t] = myE;
  bC[t] = myC;
  sm[t] = mv;
  sa[t] = ac;
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if (t < s) {
      const double e2 = bE[t + s];
      const int c2 = bC[t + s];
      const bool take = (bC[t] == INT_MAX) || (c2 != INT_MAX && (e2 < bE[t] || (e2 == bE[t] && c2 < bC[t])));
      if (take) {
        bE[t] = e2;
        bC[t] = c2;
      }
      sm[t] += sm[t + s];
      sa[t] += sa[t + s];
    }
    __syncthreads();
  }
  const int bc = bC[0];
  const int q = bc - static_cast<int>(P.chain0);
  const int be = P.ens0 + q / kLanes, bl = q % kLanes;
  const DevEns &bit = ens[be];
  const bool xyz = P.level == PMMC_LEVEL_SMOOTH;
  for (int i = t; i < 3 * P.n; i += blockDim.x)
    P.out_pos[i] = bit.pos[pos_index(xyz, i / 3, i % 3, static_cast<unsigned>(bl))];
  __syncthreads();
  const float *x = P.out_pos;
  double ret;
  if (P.level == PMMC_LEVEL_HEATMAP) {
    ret = exact_heatmap(P, x, red);
  } else if (P.level == PMMC_LEVEL_ARCS) {
    ret = *P.offset + exact_arcs_terms(P, x, red) + exact_sphere(P, x, P.sched.sphere_radius, red);
  } else {
    double bs[2];
    exact_smooth(P, x, red, bs);
    ret = bs[0] * P.wd + bs[1] * P.wa;
  }
  if (t == 0) {
    DevResult *R = P.result;
    const EnsState *st = bit.st;
    R->energy = st->E[bl];
    R->resummed = ret;
    R->best = st->best[bl];
    R->moves = sm[0];
    R->accepted = sa[0];
    R->best_moves = st->iters[bl];
    R->best_accepted = st->acc[bl];
    R->best_chain = bc;
    R->milestones = st->milestones[bl];
    R->stop_reason = st->status[bl];
    R->n_chains = P.n_chains;
  }
}

__global__ __launch_bounds__(256) void k_finalize(DevProblem *probs, const DevEns *__restrict__ ens) {
  __shared__ FinalizeLds L;
  finalize_body(probs[blockIdx.x], ens, L);
}

__global__ __launch_bounds__(256) void k_finalize_refs(const JobRef *__restrict__ refs) {
  __shared__ FinalizeLds L;
  finalize_body(*refs[blockIdx.x].prob, refs[blockIdx.x].ens, L);
}

// ------------------------------------------------------------------ //
// host side
// ------------------------------------------------------------------ //
// Ensembles whose positions fit this many bytes of LDS are staged there for
// the whole launch (one workgroup per ensemble; the LDS use bounds residency).
// Off by default: on MI355X sixteen resident wavefronts per CU reading
// positions through L2 beat three LDS-staged ones (N=64 heatmap: 5.4 vs 3.0
// G moves/s; N=24: equal).
#ifndef PMMC_STAGE_MAX_BYTES
#define PMMC_STAGE_MAX_BYTES 0
#endif
static constexpr int kStageMaxBytes = PMMC_STAGE_MAX_BYTES;

bool stage_ok(int n) { return static_cast<size_t>(n) * 3 * kLanes * sizeof(float) <= kStageMaxBytes; }

int pick_shift_of(int level, int n) {
  return (level == PMMC_LEVEL_HEATMAP && n > 32 && n <= 64 && !stage_ok(n)) ? 0 : kPickShift;
}

int resident_ensembles_per_cu(const HostProblem &h) {
  // keyed by everything that selects the kernel: level, staging, the
  // register-resident form and, for it, the table symmetry (the lane-pair
  // kernel and the one-lane-per-chain kernel differ in occupancy)
  static std::atomic<int> cache[32];
  const bool hreg = !stage_ok(h.n) && hreg_halves(h) > 0;
  const bool stage = stage_ok(h.n);
  const bool asym = hreg && !h.symmetric;
  const int key = (h.level & 3) * 8 + (hreg ? 1 : 0) + (stage ? 2 : 0) + (asym ? 4 : 0);
  const int known = cache[key].load(std::memory_order_relaxed);
  if (known > 0) return known;
  int blocks = 0;
  hipError_t e = hipErrorInvalidValue;
  if (hreg) {
    blocks = hreg_resident_per_cu(!h.symmetric);
    e = blocks > 0 ? hipSuccess : hipErrorInvalidValue;
  } else {
    const size_t lds = stage ? static_cast<size_t>(h.n) * 3 * kLanes * sizeof(float) : 0;
    switch (h.level) {
      case PMMC_LEVEL_HEATMAP:
        e = stage ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_anneal<PMMC_LEVEL_HEATMAP, true>, kLanes, lds)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_anneal<PMMC_LEVEL_HEATMAP, false>, kLanes, 0);
        break;
      case PMMC_LEVEL_ARCS:
        e = stage ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_anneal<PMMC_LEVEL_ARCS, true>, kLanes, lds)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_anneal<PMMC_LEVEL_ARCS, false>, kLanes, 0);
        break;
      default:
        e = stage ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_anneal<PMMC_LEVEL_SMOOTH, true>, kLanes, lds)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_anneal<PMMC_LEVEL_SMOOTH, false>, kLanes, 0);
    }
  }
  if (e != hipSuccess || blocks <= 0) {
    (void)hipGetLastError();
    blocks = 16;
  }
  cache[key].store(blocks, std::memory_order_relaxed);
  return blocks;
}

LaunchGroup group_of(const HostProblem &h) {
  LaunchGroup g{};
  g.level = h.level;
  g.stage = stage_ok(h.n);
  g.n = g.stage ? h.n : 0;
  g.hreg = g.stage ? 0 : hreg_halves(h);
  g.asym = g.hreg > 0 && !h.symmetric;
  return g;
}

std::tuple<int, int, int, int, int> group_key(const LaunchGroup &g) {
  return std::make_tuple(g.level, g.stage ? 1 : 0, g.n, g.hreg, g.asym ? 1 : 0);
}

template <int LEVEL>
static hipError_t launch_level(hipStream_t s, const DevProblem *probs, const DevEns *ens, int n_ens,
                               int max_blocks, bool stage, int n_for_stage) {
  if (n_ens <= 0) return hipSuccess;
  if (stage) {
    const size_t lds = static_cast<size_t>(n_for_stage) * 3 * kLanes * sizeof(float);
    hipLaunchKernelGGL((k_anneal<LEVEL, true>), dim3(n_ens), dim3(kLanes), lds, s, probs, ens,
                       max_blocks);
  } else {
    hipLaunchKernelGGL((k_anneal<LEVEL, false>), dim3(n_ens), dim3(kLanes), 0, s, probs, ens,
                       max_blocks);
  }
  return hipGetLastError();
}

// Everything a batch needs on the device.
hipError_t launch_groups(hipStream_t s, const DevProblem *probs, const DevEns *ens,
                         const std::vector<LaunchGroup> &groups, int max_blocks) {
  hipError_t err = hipSuccess;
  for (const auto &g : groups) {
    const DevEns *e = ens + g.ens0;
    switch (g.level) {
      case PMMC_LEVEL_HEATMAP:
        if (g.hreg > 0)
          err = launch_heatmap_reg(s, probs, e, g.n_ens, max_blocks, g.hreg, g.asym);
        else
          err = launch_level<PMMC_LEVEL_HEATMAP>(s, probs, e, g.n_ens, max_blocks, g.stage, g.n);
        break;
      case PMMC_LEVEL_ARCS:
        if (g.cpw == 32 && !g.stage)
          err = launch_arcs_pair(s, probs, e, g.n_ens, max_blocks);
        else
          err = launch_level<PMMC_LEVEL_ARCS>(s, probs, e, g.n_ens, max_blocks, g.stage, g.n);
        break;
      default:
        err = launch_level<PMMC_LEVEL_SMOOTH>(s, probs, e, g.n_ens, max_blocks, g.stage, g.n);
    }
    if (err != hipSuccess) return err;
  }
  return err;
}

// Workgroups per ensemble for replicating n*3*64 floats (256 per pass, at
// most 64 workgroups per ensemble), and ensembles per grid row.
static dim3 init_grid(int max_n, int n_ens) {
  const size_t total = static_cast<size_t>(3 * std::max(max_n, 1)) * kLanes;
  const unsigned gx = static_cast<unsigned>(std::min<size_t>(64, (total + 255) / 256));
  const unsigned gy = static_cast<unsigned>(std::min(std::max(n_ens, 1), 65535));
  return dim3(gx, gy);
}

hipError_t launch_init(hipStream_t s, DevProblem *probs, int n_probs, const DevEns *ens, int n_ens,
                       int max_n) {
  if (n_probs <= 0 || n_ens <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_init_problem, dim3(n_probs), dim3(256), 0, s, probs);
  hipLaunchKernelGGL(k_init_ens, init_grid(max_n, n_ens), dim3(256), 0, s, probs, ens, n_ens);
  return hipGetLastError();
}

hipError_t launch_init_refs(hipStream_t s, const JobRef *jobs, int n_jobs, const EnsRef *ens,
                            int n_ens, int max_n) {
  if (n_jobs <= 0 || n_ens <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_init_problem_refs, dim3(n_jobs), dim3(256), 0, s, jobs);
  hipLaunchKernelGGL(k_init_ens_refs, init_grid(max_n, n_ens), dim3(256), 0, s, ens, n_ens);
  return hipGetLastError();
}

hipError_t launch_finalize_refs(hipStream_t s, const JobRef *jobs, int n_jobs) {
  if (n_jobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize_refs, dim3(n_jobs), dim3(256), 0, s, jobs);
  return hipGetLastError();
}

hipError_t launch_finalize(hipStream_t s, DevProblem *probs, int n_probs, const DevEns *ens) {
  hipLaunchKernelGGL(k_finalize, dim3(n_probs), dim3(256), 0, s, probs, ens);
  return hipGetLastError();
}

// A problem's device data in two parts: the static tables the host writes
// (carved contiguously, so they go up in one copy) and the device-written
// state (initial energy, ensembles, chain queue).  The result positions
// out_pos are carved by the caller.
static void size_static(Sizer &z, const HostProblem &h) {
  z.add<int32_t>(std::max<size_t>(1, h.free_list.size()));
  z.add<float>(h.pos0.size());
  z.add<float>(h.W.size());
  z.add<float>(h.W2.size());
  z.add<float>(h.Hc.size());
  z.add<int32_t>(h.off.size());
  z.add<ArcEntry>(std::max<size_t>(1, h.ent.size()));
  z.add<int32_t>(std::max<size_t>(1, h.ej.size()));
  // coefficient tables padded by one chunk: a move loads its first chunk's
  // coefficients unconditionally (kernels.hpp, arcs_delta_fetched)
  z.add<float>(h.ec.size() + kArcsChunk);
  z.add<float>(h.eka.size() + kArcsChunk);
  z.add<float>(h.ekb.size() + kArcsChunk);
  z.add<PairEntry>(std::max<size_t>(1, h.pairs.size()));
  z.add<int32_t>(std::max<size_t>(1, h.zpairs.size()));
  z.add<float>(h.dtn.size());
}

static void size_dynamic(Sizer &z, const HostProblem &h) {
  z.add<double>(2);
  const int ne = (h.n_chains + kLanes - 1) / kLanes;
  z.add<EnsState>(ne);
  z.add<float>(static_cast<size_t>(ne) * h.pos0.size() * kLanes);
  if (h.workers > 0) {
    z.add<int32_t>(1);
    z.add<int32_t>(static_cast<size_t>(h.workers) * kLanes);
  }
}

size_t problem_bytes(const HostProblem &h) {
  Sizer z;
  size_static(z, h);
  size_dynamic(z, h);
  z.add<float>(h.pos0.size());  // out_pos
  return z.top + 256;
}

// Carves device addresses and writes the data at the same offsets of a host
// staging buffer (both regions 256-byte aligned, offsets as Sizer's).  Every
// byte it passes over is written: alignment gaps and the entries a table is
// padded to (min_count) are zeroed, so the staging buffer need not be
// cleared beforehand.
struct StageCarver {
  char *dev;
  char *host;
  size_t top = 0;
  void align() {
    const size_t at = (top + 255) & ~size_t(255);
    if (at > top) std::memset(host + top, 0, at - top);
    top = at;
  }
  template <class T>
  T *put(const std::vector<T> &v, size_t min_count = 0) {
    align();
    T *d = reinterpret_cast<T *>(dev + top);
    if (!v.empty()) std::memcpy(host + top, v.data(), sizeof(T) * v.size());
    const size_t n = std::max(v.size(), min_count);
    if (n > v.size()) std::memset(host + top + sizeof(T) * v.size(), 0, sizeof(T) * (n - v.size()));
    top += sizeof(T) * n;
    return d;
  }
};

// The context's pinned staging buffer, grown to at least `bytes`.
static pmmc_status host_stage_reserve(pmmc_ctx *ctx, size_t bytes) {
  if (bytes <= ctx->host_stage_cap) return PMMC_OK;
  const size_t cap = std::max(bytes, 2 * ctx->host_stage_cap);
  PMMC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
  ctx->host_stage = nullptr;
  ctx->host_stage_cap = 0;
  if (hipHostMalloc(reinterpret_cast<void **>(&ctx->host_stage), cap) != hipSuccess) {
    ctx->host_stage = nullptr;
    return fail(ctx, PMMC_ERR_NOMEM, "pinned staging buffer allocation failed");
  }
  ctx->host_stage_cap = cap;
  return PMMC_OK;
}

static void fill_static(StageCarver &sc, const HostProblem &h, DevProblem &P) {
  P.free_list = sc.put(h.free_list, 1);
  P.pos0 = sc.put(h.pos0);
  P.W = sc.put(h.W);
  P.W2 = h.W2.empty() ? nullptr : sc.put(h.W2);
  if (h.W2.empty()) sc.put(h.W2);
  P.Hc = sc.put(h.Hc);
  P.off = sc.put(h.off);
  P.ent = sc.put(h.ent, 1);
  P.ej = sc.put(h.ej, 1);
  P.ec = sc.put(h.ec, h.ec.size() + kArcsChunk);
  P.eka = sc.put(h.eka, h.eka.size() + kArcsChunk);
  P.ekb = sc.put(h.ekb, h.ekb.size() + kArcsChunk);
  P.pairs = sc.put(h.pairs, 1);
  P.zpairs = sc.put(h.zpairs, 1);
  P.dtn = sc.put(h.dtn);
}

// Scalars, device-written state and ensembles of one problem (problem index
// `index` in the ensembles appended to `ens`).
static void fill_dynamic(Carver &cv, const HostProblem &h, int index, DevResult *result,
                         float *out_pos, DevProblem &P, std::vector<DevEns> &ens) {
  P.level = h.level;
  P.n = h.n;
  P.nfree = static_cast<int32_t>(h.free_list.size());
  P.symmetric = h.symmetric ? 1 : 0;
  P.ldw = h.ldw > 0 ? h.ldw : h.n;
  P.npairs = static_cast<int32_t>(h.pairs.size());
  P.nnz = static_cast<int32_t>(h.ent.size());
  P.nz = static_cast<int32_t>(h.zpairs.size() / 2);
  P.ks = h.ks;
  P.kq = h.kq;
  P.ka = h.ka;
  P.wd = h.wd;
  P.wa = h.wa;
  double *e0 = cv.take<double>(2);
  P.E0 = e0;
  P.offset = e0 + 1;
  P.sched = h.sched;
  P.seed = h.seed;
  P.segment = h.segment;
  P.chain0 = h.chain0;
  P.n_chains = h.n_chains;
  P.out_pos = out_pos;
  P.result = result;
  const int ne = (h.n_chains + kLanes - 1) / kLanes;
  EnsState *sts = cv.take<EnsState>(ne);
  float *pos = cv.take<float>(static_cast<size_t>(ne) * h.pos0.size() * kLanes);
  P.workers = h.workers;
  P.pick_shift = h.pick_shift;
  if (h.workers > 0) {
    P.queue = cv.take<int32_t>(1);
    P.slot_chain = cv.take<int32_t>(static_cast<size_t>(h.workers) * kLanes);
  }
  P.ens0 = static_cast<int32_t>(ens.size());
  P.n_ens = ne;
  for (int e = 0; e < ne; ++e) {
    DevEns d;
    d.problem = index;
    d.chain0 = h.chain0 + static_cast<uint32_t>(e * kLanes);
    d.nlanes = std::min(kLanes, h.n_chains - e * kLanes);
    d.pad = 0;
    d.pos = pos + static_cast<size_t>(e) * h.pos0.size() * kLanes;
    d.st = sts + e;
    ens.push_back(d);
  }
}

// Upload one problem into carved device memory (its static tables staged in
// `stage`, which must stay alive until the stream has synchronised); append
// its ensembles (with problem index `index`) to `ens`.
void fill_problem(Carver &cv, hipStream_t s, const HostProblem &h, int index, DevResult *result,
                  DevProblem &P, std::vector<DevEns> &ens, hipError_t &err, std::vector<char> &stage) {
  std::memset(&P, 0, sizeof(P));
  Sizer zs;
  size_static(zs, h);
  stage.assign(zs.top, 0);
  char *sdev = cv.take<char>(zs.top);
  StageCarver sc{sdev, stage.data()};
  fill_static(sc, h, P);
  if (err == hipSuccess && zs.top > 0)
    err = hipMemcpyAsync(sdev, stage.data(), zs.top, hipMemcpyHostToDevice, s);
  float *out_pos = cv.take<float>(h.pos0.size());
  fill_dynamic(cv, h, index, result, out_pos, P, ens);
}

// Bytes of a problem's static-table block and of its result positions,
// each block starting 256-byte aligned (so a region's size does not depend
// on the order of its problems).
static size_t static_block(const HostProblem &h) {
  Sizer z;
  size_static(z, h);
  return (z.top + 255) & ~size_t(255);
}

static size_t out_block(const HostProblem &h) {
  return (sizeof(float) * h.pos0.size() + 255) & ~size_t(255);
}

// Device bytes batch_setup carves for these problems, in its order: problem
// and result arrays, counters, ensembles, every problem's static tables,
// every problem's result positions, every problem's device-written state.
size_t batch_bytes(const std::vector<HostProblem> &hp) {
  Sizer z;
  z.add<DevProblem>(hp.size());
  z.add<DevResult>(hp.size());
  z.add<int>(64);
  size_t n_ens = 0, sb = 0, ob = 0;
  for (auto &h : hp) {
    n_ens += (h.n_chains + kLanes - 1) / kLanes;
    sb += static_block(h);
    ob += out_block(h);
  }
  z.add<DevEns>(n_ens);
  z.add<char>(sb);
  z.add<char>(ob);
  for (auto &h : hp) size_dynamic(z, h);
  return z.top + 4096;
}

// Carves the batch from `span` when it has a base (the caller's device
// memory, nothing is allocated), else from the context's arena.  The static
// tables of every problem go up in one copy.
pmmc_status batch_setup(pmmc_ctx *ctx, std::vector<HostProblem> &hp, BatchDev &bd, DevSpan span) {
  size_t n_ens = 0, sb = 0, ob = 0;
  for (auto &h : hp) {
    n_ens += (h.n_chains + kLanes - 1) / kLanes;
    sb += static_block(h);
    ob += out_block(h);
  }
  const size_t need = batch_bytes(hp);
  // a caller's span (the allocation-free entry point) allocates nothing: its
  // tables are staged in pageable memory; otherwise in the context's pinned
  // buffer, uploaded by DMA without a runtime copy
  const bool caller_span = span.base != nullptr;
  if (span.base) {
    if (need > span.cap)
      return fail(ctx, PMMC_ERR_NOMEM, "the problem needs " + std::to_string(need) +
                                           " bytes of device memory; the span holds " +
                                           std::to_string(span.cap));
  } else {
    pmmc_status ps = arena_reserve(ctx, need);
    if (ps != PMMC_OK) return ps;
    span.base = ctx->arena;
    span.cap = ctx->arena_cap;
  }
  Carver cv{span.base, 0, span.cap};
  hipStream_t s = ctx->stream;
  hipError_t err = hipSuccess;
  bd.probs = cv.take<DevProblem>(hp.size());
  bd.results = cv.take<DevResult>(hp.size());
  bd.counter = cv.take<int>(64);
  bd.ens = cv.take<DevEns>(n_ens);
  char *sdev = cv.take<char>(sb);
  bd.outs = cv.take<char>(ob);
  bd.outs_bytes = ob;
  char *host = nullptr;
  if (caller_span) {
    bd.stage.resize(sb);
    host = bd.stage.data();
  } else {
    pmmc_status ps = host_stage_reserve(ctx, sb);
    if (ps != PMMC_OK) return ps;
    host = ctx->host_stage;
  }
  StageCarver sc{sdev, host};
  size_t otop = 0;
  bd.out_off.assign(hp.size(), 0);
  bd.hprobs.resize(hp.size());
  bd.hens.clear();
  bd.groups.clear();
  // order ensembles by launch group so each group is contiguous, and within
  // a group the largest problems first: the dispatcher hands the first
  // workgroups of a launch to the CUs one per SIMD before it doubles up, so
  // the longest-running ensembles get a SIMD each and share it only with
  // small problems' ensembles, which stop early.  Results do not depend on
  // where an ensemble sits in the grid.
  std::vector<int> order(hp.size());
  for (size_t i = 0; i < hp.size(); ++i) order[i] = static_cast<int>(i);
  auto key = [&](int i) { return group_key(group_of(hp[i])); };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const auto ka = key(a), kb = key(b);
    return ka < kb || (ka == kb && hp[a].n > hp[b].n);
  });
  for (int oi : order) {
    HostProblem &h = hp[oi];
    const LaunchGroup g = group_of(h);
    if (bd.groups.empty() || group_key(bd.groups.back()) != group_key(g)) {
      bd.groups.push_back(g);
      bd.groups.back().ens0 = static_cast<int>(bd.hens.size());
    }
    const size_t before = bd.hens.size();
    DevProblem &P = bd.hprobs[oi];
    std::memset(&P, 0, sizeof(P));
    sc.align();
    fill_static(sc, h, P);
    bd.out_off[oi] = otop;
    float *out_pos = reinterpret_cast<float *>(bd.outs + otop);
    otop += out_block(h);
    fill_dynamic(cv, h, oi, bd.results + oi, out_pos, P, bd.hens);
    bd.groups.back().n_ens += static_cast<int>(bd.hens.size() - before);
  }
  bd.n_ens_total = static_cast<int>(bd.hens.size());
  if (sc.top < sb) std::memset(host + sc.top, 0, sb - sc.top);
  // one upload of every problem's static tables, then the descriptors
  if (sb > 0) err = hipMemcpyAsync(sdev, host, sb, hipMemcpyHostToDevice, s);
  if (err == hipSuccess)
    err = hipMemcpyAsync(bd.probs, bd.hprobs.data(), sizeof(DevProblem) * hp.size(),
                         hipMemcpyHostToDevice, s);
  if (err == hipSuccess)
    err = hipMemcpyAsync(bd.ens, bd.hens.data(), sizeof(DevEns) * bd.hens.size(),
                         hipMemcpyHostToDevice, s);
  if (err != hipSuccess) return fail(ctx, PMMC_ERR_HIP, std::string("upload: ") + hipGetErrorString(err));
  int max_n = 1;
  for (auto &h : hp) max_n = std::max(max_n, h.n);
  err = launch_init(s, bd.probs, static_cast<int>(hp.size()), bd.ens, bd.n_ens_total, max_n);
  if (err != hipSuccess) return fail(ctx, PMMC_ERR_HIP, std::string("init: ") + hipGetErrorString(err));
  return PMMC_OK;
}

hipError_t batch_step(pmmc_ctx *ctx, const BatchDev &bd, int max_blocks) {
  return launch_groups(ctx->stream, bd.probs, bd.ens, bd.groups, max_blocks);
}

hipError_t batch_count_active(pmmc_ctx *ctx, const BatchDev &bd, int *host_count) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, ctx->stream, bd.counter);
  const int grid = std::min(bd.n_ens_total, 1024);
  hipLaunchKernelGGL(k_count_active, dim3(grid), dim3(64), 0, ctx->stream, bd.ens, bd.n_ens_total, bd.counter);
  hipError_t err = hipMemcpyAsync(host_count, bd.counter, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(ctx->stream);
}

pmmc_status batch_finish(pmmc_ctx *ctx, std::vector<HostProblem> &hp, BatchDev &bd,
                         std::vector<float *> &outs, std::vector<pmmc_anneal_stats *> &stats,
                         double kernel_ms) {
  hipStream_t s = ctx->stream;
  hipError_t err = launch_finalize(s, bd.probs, static_cast<int>(hp.size()), bd.ens);
  std::vector<DevResult> res(hp.size());
  if (err == hipSuccess)
    err = hipMemcpyAsync(res.data(), bd.results, sizeof(DevResult) * hp.size(), hipMemcpyDeviceToHost, s);
  // every problem's result positions come back in one copy
  std::vector<char> tmp(bd.outs_bytes);
  if (err == hipSuccess && bd.outs_bytes > 0)
    err = hipMemcpyAsync(tmp.data(), bd.outs, bd.outs_bytes, hipMemcpyDeviceToHost, s);
  if (err == hipSuccess) err = hipStreamSynchronize(s);
  if (err != hipSuccess) return fail(ctx, PMMC_ERR_HIP, std::string("finalize: ") + hipGetErrorString(err));
  for (size_t i = 0; i < hp.size(); ++i) {
    const float *res_pos = reinterpret_cast<const float *>(tmp.data() + bd.out_off[i]);
    if (outs[i]) {
      if (hp[i].level == PMMC_LEVEL_SMOOTH) {
        // only non-fixed beads are written back (ParallelMonteCarloSmooth.cu:485-491)
        for (int b = 0; b < hp[i].n; ++b)
          if (!hp[i].fixed[b])
            for (int k = 0; k < 3; ++k) outs[i][3 * b + k] = res_pos[3 * b + k];
      } else {
        std::memcpy(outs[i], res_pos, sizeof(float) * hp[i].pos0.size());
      }
    }
    if (stats[i]) {
      pmmc_anneal_stats &S = *stats[i];
      const DevResult &R = res[i];
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
      S.kernel_ms = kernel_ms;
    }
  }
  return PMMC_OK;
}

// Moves summed over every chain of the batch (synchronises the stream).
static hipError_t batch_moves(pmmc_ctx *ctx, const BatchDev &bd, unsigned long long *host) {
  unsigned long long *d = reinterpret_cast<unsigned long long *>(bd.counter + 2);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, d);
  const int grid = std::max(1, std::min(bd.n_ens_total, 1024));
  hipLaunchKernelGGL(k_sum_iters, dim3(grid), dim3(64), 0, ctx->stream, bd.ens, bd.n_ens_total, d);
  hipError_t err = hipMemcpyAsync(host, d, sizeof(*host), hipMemcpyDeviceToHost, ctx->stream);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(ctx->stream);
}

static void set_cpw(BatchDev &bd, int cpw) {
  for (auto &g : bd.groups)
    if (g.level == PMMC_LEVEL_ARCS) g.cpw = cpw;
}

// Launch-shape tuner: fields of `sh` that are 0 are chosen by timing short
// steps of the run itself (the chains advance; nothing is repeated).
static pmmc_status tune_shape(pmmc_ctx *ctx, BatchDev &bd, BatchShape &sh, int max_block) {
  hipStream_t s = ctx->stream;
  const bool has_arcs = std::any_of(bd.groups.begin(), bd.groups.end(),
                                    [](const LaunchGroup &g) { return g.level == PMMC_LEVEL_ARCS && !g.stage; });
  std::vector<int> cands;
  if (sh.cpw > 0 || !has_arcs) cands.push_back(sh.cpw > 0 ? sh.cpw : 64);
  else cands = {64, 32};
  // a probe step of about 1,024 moves per chain
  const int probe = std::max(1, 1024 / std::max(1, max_block));
  double best_rate = -1.0, best_ms = 0.0;
  int best_cpw = cands[0];
  unsigned long long before = 0, after = 0;
  hipError_t err = batch_moves(ctx, bd, &before);
  for (int cpw : cands) {
    if (err != hipSuccess) break;
    set_cpw(bd, cpw);
    PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
    err = batch_step(ctx, bd, probe);
    if (err != hipSuccess) break;
    PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev2, s));
    err = batch_moves(ctx, bd, &after);
    if (err != hipSuccess) break;
    float ms = 0.f;
    PMMC_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev1, ctx->ev2));
    const double rate = ms > 0.f ? static_cast<double>(after - before) / ms : 0.0;
    if (rate > best_rate) {
      best_rate = rate;
      best_cpw = cpw;
      best_ms = ms;
    }
    before = after;
  }
  if (err != hipSuccess) return fail(ctx, PMMC_ERR_HIP, std::string("tune: ") + hipGetErrorString(err));
  if (sh.cpw <= 0) sh.cpw = best_cpw;
  set_cpw(bd, sh.cpw);
  const double per_block_ms = best_ms > 0.0 ? best_ms / probe : 0.0;
  if (sh.bps <= 0)
    sh.bps = per_block_ms > 0.0 ? static_cast<int>(std::min(1e6, std::max(1.0, 20.0 / per_block_ms)))
                                : std::max(1, 8192 / std::max(1, max_block));
  if (sh.spg <= 0)
    sh.spg = per_block_ms > 0.0
                 ? static_cast<int>(std::min(64.0, std::max(1.0, 100.0 / (per_block_ms * sh.bps))))
                 : 8;
  sh.tuned = 1;
  return PMMC_OK;
}

// Run a set of prepared problems to completion.  sh.bps <= 0 with sh.tune
// off: a single launch per level group that runs every chain to its end.
// Otherwise step launches of sh.bps blocks, sh.spg of them captured per
// hipGraph replay when sh.graph, with launch-shape fields left 0 autotuned
// when sh.tune.
pmmc_status run_problems(pmmc_ctx *ctx, std::vector<HostProblem> &hp, std::vector<float *> &outs,
                         std::vector<pmmc_anneal_stats *> &stats, BatchShape &sh, DevSpan span) {
  BatchDev bd;
  pmmc_status ps = batch_setup(ctx, hp, bd, span);
  if (ps != PMMC_OK) return ps;
  hipStream_t s = ctx->stream;
  if (sh.cpw > 0) set_cpw(bd, sh.cpw);
  if (!sh.tune && sh.bps <= 0) {
    PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
    hipError_t err = batch_step(ctx, bd, INT_MAX);
    if (err != hipSuccess) return fail(ctx, PMMC_ERR_HIP, std::string("anneal: ") + hipGetErrorString(err));
  } else {
    int max_block = 1;
    for (const auto &h : hp) max_block = std::max(max_block, h.sched.block);
    // the tuner's probe steps are part of the timed run
    PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
    if (sh.tune) {
      ps = tune_shape(ctx, bd, sh, max_block);
      if (ps != PMMC_OK) return ps;
    }
    // the captured step loop; destroyed on every return path
    struct GraphGuard {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      ~GraphGuard() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
      }
    } gg;
    const int bps = sh.bps > 0 ? sh.bps : 1;
    if (sh.graph) {
      if (sh.spg <= 0) sh.spg = 8;
      PMMC_HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      hipError_t cerr = hipSuccess;
      for (int i = 0; i < sh.spg && cerr == hipSuccess; ++i) cerr = batch_step(ctx, bd, bps);
      hipError_t eerr = hipStreamEndCapture(s, &gg.graph);
      if (cerr != hipSuccess || eerr != hipSuccess)
        return fail(ctx, PMMC_ERR_HIP, "graph capture failed");
      PMMC_HIP_TRY(ctx, hipGraphInstantiate(&gg.exec, gg.graph, nullptr, nullptr, 0));
    }
    for (;;) {
      int active = 0;
      hipError_t err = batch_count_active(ctx, bd, &active);
      if (err != hipSuccess) return fail(ctx, PMMC_ERR_HIP, std::string("count: ") + hipGetErrorString(err));
      if (active == 0) break;
      if (sh.graph) err = hipGraphLaunch(gg.exec, s);
      else err = batch_step(ctx, bd, bps);
      if (err != hipSuccess) return fail(ctx, PMMC_ERR_HIP, std::string("step: ") + hipGetErrorString(err));
    }
  }
  PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
  PMMC_HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  return batch_finish(ctx, hp, bd, outs, stats, ms);
}

}  // namespace pmmc

// ------------------------------------------------------------------ //
// C ABI
// ------------------------------------------------------------------ //
using namespace pmmc;

static pmmc_status plain(pmmc_ctx *ctx, int level, const void *prob, const pmmc_schedule *s,
                         const pmmc_run *run, float *pos_out, pmmc_anneal_stats *stats) {
  if (!ctx || !prob || !s || !run || !pos_out) return fail(ctx, PMMC_ERR_INVALID, "null argument");
  std::vector<HostProblem> hp(1);
  pmmc_status ps = prepare_problem(ctx, level, prob, *s, *run, hp[0]);
  if (ps != PMMC_OK) return ps;
  std::vector<float *> outs{pos_out};
  std::vector<pmmc_anneal_stats *> st{stats};
  BatchShape sh;
  return run_problems(ctx, hp, outs, st, sh);
}

extern "C" pmmc_status pmmc_heatmap_anneal(pmmc_ctx *ctx, const pmmc_heatmap_problem *p,
                                           const pmmc_schedule *s, const pmmc_run *run,
                                           float *pos_out, pmmc_anneal_stats *stats) {
  return plain(ctx, PMMC_LEVEL_HEATMAP, p, s, run, pos_out, stats);
}

extern "C" pmmc_status pmmc_arcs_anneal(pmmc_ctx *ctx, const pmmc_arcs_problem *p,
                                        const pmmc_schedule *s, const pmmc_run *run,
                                        float *pos_out, pmmc_anneal_stats *stats) {
  return plain(ctx, PMMC_LEVEL_ARCS, p, s, run, pos_out, stats);
}

extern "C" pmmc_status pmmc_smooth_anneal(pmmc_ctx *ctx, const pmmc_smooth_problem *p,
                                          const pmmc_schedule *s, const pmmc_run *run,
                                          float *pos_out, pmmc_anneal_stats *stats) {
  return plain(ctx, PMMC_LEVEL_SMOOTH, p, s, run, pos_out, stats);
}

namespace pmmc {

// Jobs are prepared on host threads: building the tables (an n x n scan per
// arc or heatmap problem) is most of a large batch's host time.  Each thread
// reports errors through its own copy of the context.  Jobs that repeat an
// earlier job's (level, problem) copy its tables in a second pass; a job
// whose representative failed prepares alone, so every job gets the error
// prepare_problem would give it.
pmmc_status prepare_jobs(pmmc_ctx *ctx, const pmmc_job *jobs, int n_jobs, std::vector<HostProblem> &hp,
                         int *failed) {
  hp.assign(n_jobs, HostProblem{});
  std::vector<int> rep(n_jobs);
  std::map<std::pair<int, const void *>, int> first;
  for (int i = 0; i < n_jobs; ++i)
    rep[i] = first.emplace(std::make_pair(jobs[i].level, jobs[i].problem), i).first->second;
  std::vector<pmmc_status> pst(n_jobs, PMMC_OK);
  std::vector<std::string> pmsg(n_jobs);
  const int nt = std::max(1, std::min({n_jobs / 4, 16, static_cast<int>(std::thread::hardware_concurrency())}));
  auto pass = [&](bool copies) {
    auto body = [&](int t) {
      pmmc_ctx local;
      local.device = ctx->device;
      local.cu_count = ctx->cu_count;
      local.hbm_bytes = ctx->hbm_bytes;
      local.default_seed = ctx->default_seed;
      for (int i = t; i < n_jobs; i += nt) {
        if ((rep[i] != i) != copies) continue;
        const pmmc_job &j = jobs[i];
        if (copies && pst[rep[i]] == PMMC_OK) {
          try {
            hp[i] = hp[rep[i]];
            pst[i] = set_run(&local, j.schedule, j.run, hp[i]);
          } catch (const std::bad_alloc &) {
            hp[i] = HostProblem{};
            pst[i] = fail(&local, PMMC_ERR_NOMEM, "host memory exhausted while copying the job's tables");
          }
        } else {
          pst[i] = prepare_problem(&local, j.level, j.problem, j.schedule, j.run, hp[i]);
        }
        if (pst[i] != PMMC_OK) pmsg[i] = local.last_error;
      }
    };
    if (nt == 1) {
      body(0);
    } else {
      std::vector<std::thread> th;
      for (int t = 0; t < nt; ++t) th.emplace_back(body, t);
      for (auto &x : th) x.join();
    }
  };
  pass(false);
  pass(true);
  for (int i = 0; i < n_jobs; ++i)
    if (pst[i] != PMMC_OK) {
      if (failed) *failed = i;
      return fail(ctx, pst[i], pmsg[i]);
    }
  return PMMC_OK;
}

}  // namespace pmmc

static pmmc_status anneal_batch(pmmc_ctx *ctx, const pmmc_job *jobs, int32_t n_jobs,
                                pmmc_batch_opts *opts) {
  if (!ctx || (!jobs && n_jobs > 0) || n_jobs < 0) return fail(ctx, PMMC_ERR_INVALID, "bad batch");
  if (n_jobs == 0) return PMMC_OK;
  std::vector<HostProblem> hp;
  std::vector<float *> outs(n_jobs);
  std::vector<pmmc_anneal_stats *> st(n_jobs);
  for (int i = 0; i < n_jobs; ++i) {
    if (!jobs[i].pos_out) return fail(ctx, PMMC_ERR_INVALID, "job without pos_out");
    outs[i] = jobs[i].pos_out;
    st[i] = jobs[i].stats;
  }
  int bad = -1;
  const pmmc_status pp = prepare_jobs(ctx, jobs, n_jobs, hp, &bad);
  if (pp != PMMC_OK) return fail(ctx, pp, "job " + std::to_string(bad) + ": " + ctx->last_error);
  BatchShape sh;
  sh.bps = opts ? opts->blocks_per_step : 0;
  sh.spg = opts ? opts->steps_per_graph : 0;
  sh.graph = opts ? opts->use_graph != 0 : false;
  sh.cpw = opts ? opts->chains_per_wave : 0;
  if (sh.cpw != 0 && sh.cpw != 64 && sh.cpw != 32)
    return fail(ctx, PMMC_ERR_INVALID, "chains_per_wave must be 0, 32 or 64");
  if (sh.bps < -1) return fail(ctx, PMMC_ERR_INVALID, "blocks_per_step must be >= -1");
  const bool one_launch = sh.bps == -1;
  if (one_launch) {
    // one launch per level group that runs every chain to its end: nothing
    // to step, capture or tune but the chains per wavefront (64 if unset)
    sh.bps = 0;
    sh.graph = false;
    sh.tune = false;
    if (sh.cpw <= 0) sh.cpw = 64;
  } else {
    sh.tune = sh.bps <= 0 || sh.cpw <= 0 || (sh.graph && sh.spg <= 0);
  }
  const pmmc_status ps = run_problems(ctx, hp, outs, st, sh);
  if (ps == PMMC_OK && opts) {
    opts->blocks_per_step = one_launch ? -1 : sh.bps;
    opts->steps_per_graph = sh.graph ? sh.spg : 0;
    opts->chains_per_wave = sh.cpw > 0 ? sh.cpw : 64;
    opts->tuned = sh.tuned;
  }
  return ps;
}

extern "C" pmmc_status pmmc_anneal_batch(pmmc_ctx *ctx, const pmmc_job *jobs, int32_t n_jobs,
                                         pmmc_batch_opts *opts) {
  return no_throw(ctx, [&] { return anneal_batch(ctx, jobs, n_jobs, opts); });
}

// ------------------------------------------------------------------ //
// per-move ΔE on injected proposals (test hook): the same device functions
// as k_anneal, evaluated against one fixed state
// ------------------------------------------------------------------ //
namespace pmmc {

template <int LEVEL>
__global__ __launch_bounds__(64) void k_delta(const DevProblem *__restrict__ probs,
                                              const DevEns *__restrict__ ens, int n_prop,
                                              const int32_t *__restrict__ beads,
                                              const float *__restrict__ dsp, float *__restrict__ out) {
  // one proposal per wavefront: the bead is wave-uniform, exactly as on the
  // annealing path (all 64 lanes hold the same replicated state)
  const DevProblem &P = probs[0];
  const int lane = threadIdx.x;
  const int i = blockIdx.x;
  if (i >= n_prop) return;
  const GlobalViewT<LEVEL> X = level_view<LEVEL>(ens[0].pos, lane);
  const int n = P.n;
  const int p = __builtin_amdgcn_readfirstlane(beads[i]);
  const float px = X.get(p, 0), py = X.get(p, 1), pz = X.get(p, 2);
  const float nx = move_coord(px, dsp[3 * i]), ny = move_coord(py, dsp[3 * i + 1]);
  const float nz = move_coord(pz, dsp[3 * i + 2]);
  float dE;
  if constexpr (LEVEL == PMMC_LEVEL_HEATMAP) {
    dE = heatmap_delta(P, X, p, px, py, pz, nx, ny, nz);
  } else if constexpr (LEVEL == PMMC_LEVEL_ARCS) {
    dE = arcs_delta(P, X, p, px, py, pz, nx, ny, nz);
  } else {
    float qx[5], qy[5], qz[5];
    for (int o = 0; o < 5; ++o) {
      const int idx = p - 2 + o;
      const bool in = (o != 2) && idx >= 0 && idx < n;
      qx[o] = in ? X.get(idx, 0) : 0.f;
      qy[o] = in ? X.get(idx, 1) : 0.f;
      qz[o] = in ? X.get(idx, 2) : 0.f;
    }
    const float d0m = p > 0 ? fmaxf(P.dtn[p - 1], 1e-6f) : 1.f;
    const float d0p = p < n - 1 ? fmaxf(P.dtn[p], 1e-6f) : 1.f;
    dE = smooth_delta(P, p, qx, qy, qz, px, py, pz, nx, ny, nz, d0m, d0p);
  }
  const float r = P.sched.sphere_radius;
  if (r > 0.f) dE += sphere_penalty(nx, ny, nz, r) - sphere_penalty(px, py, pz, r);
  if (lane == 0) out[i] = dE;
}

}  // namespace pmmc

extern "C" pmmc_status pmmc_delta_energy(pmmc_ctx *ctx, int level, const void *problem,
                                         float sphere_radius, int32_t n_prop, const int32_t *beads,
                                         const float *disp, float *delta_out) {
  if (!ctx || !problem || n_prop < 0 || (n_prop > 0 && (!beads || !disp || !delta_out)))
    return fail(ctx, PMMC_ERR_INVALID, "delta: bad arguments");
  if (n_prop == 0) return PMMC_OK;
  pmmc_schedule s;
  std::memset(&s, 0, sizeof(s));
  s.T0 = 1.0;
  s.dt = 1.0;
  s.block = 1;
  s.milestone_period = 1;
  s.sphere_radius = sphere_radius;
  s.guard_energy = -INFINITY;
  pmmc_run run{1, 0, 0, kLanes};
  std::vector<HostProblem> hp(1);
  pmmc_status ps = prepare_problem(ctx, level, problem, s, run, hp[0]);
  if (ps != PMMC_OK) return ps;
  for (int i = 0; i < n_prop; ++i)
    if (beads[i] < 0 || beads[i] >= hp[0].n) return fail(ctx, PMMC_ERR_INVALID, "delta: bead out of range");
  BatchDev bd;
  ps = batch_setup(ctx, hp, bd, DevSpan{});
  if (ps != PMMC_OK) return ps;
  hipStream_t st = ctx->stream;
  // proposals live after the batch in a second span of the arena; grow it by
  // re-running setup if needed
  const size_t need = (sizeof(int32_t) + 4 * sizeof(float)) * static_cast<size_t>(n_prop) + 4096;
  int32_t *d_b = nullptr;
  float *d_d = nullptr, *d_o = nullptr;
  PMMC_HIP_TRY(ctx, hipMallocAsync(reinterpret_cast<void **>(&d_b), need, st));
  d_d = reinterpret_cast<float *>(reinterpret_cast<char *>(d_b) + ((sizeof(int32_t) * n_prop + 255) & ~size_t(255)));
  d_o = d_d + ((3 * static_cast<size_t>(n_prop) + 63) & ~size_t(63));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_b, beads, sizeof(int32_t) * n_prop, hipMemcpyHostToDevice, st));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_d, disp, sizeof(float) * 3 * n_prop, hipMemcpyHostToDevice, st));
  const int grid = n_prop;
  switch (level) {
    case PMMC_LEVEL_HEATMAP:
      hipLaunchKernelGGL(k_delta<PMMC_LEVEL_HEATMAP>, dim3(grid), dim3(kLanes), 0, st, bd.probs, bd.ens, n_prop, d_b, d_d, d_o);
      break;
    case PMMC_LEVEL_ARCS:
      hipLaunchKernelGGL(k_delta<PMMC_LEVEL_ARCS>, dim3(grid), dim3(kLanes), 0, st, bd.probs, bd.ens, n_prop, d_b, d_d, d_o);
      break;
    default:
      hipLaunchKernelGGL(k_delta<PMMC_LEVEL_SMOOTH>, dim3(grid), dim3(kLanes), 0, st, bd.probs, bd.ens, n_prop, d_b, d_d, d_o);
  }
  PMMC_HIP_TRY(ctx, hipGetLastError());
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(delta_out, d_o, sizeof(float) * n_prop, hipMemcpyDeviceToHost, st));
  PMMC_HIP_TRY(ctx, hipFreeAsync(d_b, st));
  PMMC_HIP_TRY(ctx, hipStreamSynchronize(st));
  return PMMC_OK;
}
