// Register-resident heatmap annealer for heatmaps of at most 64 beads.
//
// Same Markov chains, random streams, schedule and stop rules as
// k_anneal<PMMC_LEVEL_HEATMAP> (engine.hip); only where the positions live
// differs.  Each lane keeps its chain's whole structure in VGPRs (NB = 64
// beads per axis), so the O(N) partner loop of a move reads no memory: it
// is fully unrolled over constant register indices.  The weight table (16 KB
// for 64 beads) is copied to LDS once per launch; a move reads row p with
// uniform-address ds_read_b128.
//
// The moved bead p is wave-uniform (shared pick stream).  Indexing the
// register arrays with it -- a 64-way switch, or vector halves with indexed
// register moves -- made the compiler spill hundreds of registers, and so did
// the O(N^2) milestone re-sum inlined into the kernel.  So: the SoA copy of
// the positions in memory is kept current (an accepted move stores its bead
// there too), the moved bead's old position is read from it one move ahead
// (the pick stream does not depend on the chains' state; the prefetched copy
// is patched when the current move changes that bead), the register copy is
// updated by one select per register, and the re-sum is an out-of-line call
// that reads the memory copy.  Symmetric tables: no spills in the move loop.
//
// Two forms: k_anneal_hreg (one lane per chain, 2 waves per SIMD) runs
// asymmetric tables; k_anneal_hpair (a chain split over a lane pair by bead
// parity, 3 waves per SIMD) runs symmetric ones, 11% faster on 64 beads.
#include <hip/hip_runtime.h>

#include "ctx.hpp"
#include "engine_host.hpp"
#include "kernels.hpp"

namespace pmmc {

// ΔE of moving p from (px,py,pz) to (nx,ny,nz): exact change of the
// double-counted full score, as heatmap_delta (kernels.hpp), over registers.
// The weight rows come from LDS (row p of W at lw + p*NB; of W2 at
// lw2 + p*NB), four partners per ds_read_b128, the same address in every
// lane.
template <int NB, bool ASYM>
__device__ __forceinline__ float hreg_delta(const float *lw, const float *lw2, const float (&X)[NB],
                                            const float (&Y)[NB], const float (&Z)[NB], int p,
                                            float px, float py, float pz, float nx, float ny,
                                            float nz) {
  const float4 *w1 = reinterpret_cast<const float4 *>(lw + p * NB);
  const float4 *w2 = reinterpret_cast<const float4 *>(lw2 + p * NB);
  float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
  // One pass: both distances of a partner serve the W terms (acc0 / acc1 by
  // partner parity) and, for asymmetric tables, the W2 term (acc2, every
  // partner in increasing order); each sum keeps its own order.
#pragma unroll
  for (int q = 0; q < NB; q += 4) {
    const float4 w = w1[q / 4];
    const float wq[4] = {w.x, w.y, w.z, w.w};
    float wq2[4] = {0.f, 0.f, 0.f, 0.f};
    if constexpr (ASYM) {
      const float4 v = w2[q / 4];
      wq2[0] = v.x;
      wq2[1] = v.y;
      wq2[2] = v.z;
      wq2[3] = v.w;
    }
#pragma unroll
    for (int u = 0; u < 4; u += 2) {
      const int a = q + u, b = q + u + 1;
      const float da0 = dist3(px, py, pz, X[a], Y[a], Z[a]);
      const float da1 = dist3(nx, ny, nz, X[a], Y[a], Z[a]);
      const float db0 = dist3(px, py, pz, X[b], Y[b], Z[b]);
      const float db1 = dist3(nx, ny, nz, X[b], Y[b], Z[b]);
      const float a0 = fmaf(da0, wq[u], -1.f), a1 = fmaf(da1, wq[u], -1.f);
      const float b0 = fmaf(db0, wq[u + 1], -1.f), b1 = fmaf(db1, wq[u + 1], -1.f);
      acc0 = fmaf(a1 - a0, a1 + a0, acc0);
      acc1 = fmaf(b1 - b0, b1 + b0, acc1);
      if constexpr (ASYM) {
        const float c0 = fmaf(da0, wq2[u], -1.f), c1 = fmaf(da1, wq2[u], -1.f);
        acc2 = fmaf(c1 - c0, c1 + c0, acc2);
        const float e0 = fmaf(db0, wq2[u + 1], -1.f), e1 = fmaf(db1, wq2[u + 1], -1.f);
        acc2 = fmaf(e1 - e0, e1 + e0, acc2);
      }
    }
  }
  const float d = acc0 + acc1;
  if constexpr (!ASYM) {
    return 2.f * d;
  } else {
    return d + acc2;
  }
}

// Milestone re-sum over the memory copy of the positions, out of line so that
// its loops do not share a register allocation with the move loop.
__device__ __noinline__ double hreg_resum(const DevProblem &P, float *pos, int lane, float r) {
  const PosView V = global_view(pos, lane);
  return resum_heatmap(P, V, r);
}

// Occupancy target: 2 waves per SIMD (192 position VGPRs of 256).
template <int NB, bool ASYM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_anneal_hreg(const DevProblem *__restrict__ probs,
                                                    const DevEns *__restrict__ ens, int max_blocks) {
  extern __shared__ float lw[];  // W (and W2), NB*NB floats each
  const DevEns it = ens[blockIdx.x];
  const DevProblem &P = probs[it.problem];
  const int lane = threadIdx.x;
  EnsState *__restrict__ st = it.st;
  int status = st->status[lane];
  if (__ballot(status == 0) == 0) return;

  const int n = P.n;
  constexpr int L = NB;  // == P.ldw (rows padded with zero weights)
  float *lw2 = lw + L * L;
  // W and W2 hold n rows of L floats; rows n..L-1 of the LDS tables are zero
  for (int i = lane; i < L * L; i += kLanes) {
    const bool in = i < n * L;
    lw[i] = in ? P.W[i] : 0.f;
    if constexpr (ASYM) lw2[i] = in ? P.W2[i] : 0.f;
  }
  __syncthreads();
  // global, not generic, address space: flat loads and stores would also
  // count on lgkmcnt, so every wait for a weight-row LDS read would wait for
  // the memory copy's traffic too
  float PMMC_AS_GLOBAL *g = (float PMMC_AS_GLOBAL *)(it.pos) + lane;
  float X[NB], Y[NB], Z[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    X[j] = j < n ? g[(3 * j + 0) * kLanes] : 0.f;
    Y[j] = j < n ? g[(3 * j + 1) * kLanes] : 0.f;
    Z[j] = j < n ? g[(3 * j + 2) * kLanes] : 0.f;
  }

  const pmmc_schedule S = P.sched;
  const uint64_t seed = P.seed;
  const uint32_t segment = P.segment;
  const int nfree = P.nfree;
  const float r = S.sphere_radius;

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
    // The moved bead's old position comes from the SoA copy in memory (kept
    // current on every accepted move), read one move ahead: the pick stream
    // does not depend on the chains' state.
    U4 rbn = draw(seed, m0 >> 2, group, segment, PMMC_LEVEL_HEATMAP, kSubBead);
    int pn = cst(P.free_list)[pick(pick_word(rbn, m0), static_cast<uint32_t>(nfree))];
    float qx = g[(3 * pn + 0) * kLanes], qy = g[(3 * pn + 1) * kLanes], qz = g[(3 * pn + 2) * kLanes];
    for (int k = 0; k < S.block; ++k) {
      const uint32_t m = m0 + static_cast<uint32_t>(k);
      const int p = pn;
      const float px = qx, py = qy, pz = qz;
      const bool more = k + 1 < S.block;
      if (more) {
        const uint32_t mn = m + 1;
        if ((mn & 3u) == 0) rbn = draw(seed, mn >> 2, group, segment, PMMC_LEVEL_HEATMAP, kSubBead);
        pn = cst(P.free_list)[pick(pick_word(rbn, mn), static_cast<uint32_t>(nfree))];
        qx = g[(3 * pn + 0) * kLanes];
        qy = g[(3 * pn + 1) * kLanes];
        qz = g[(3 * pn + 2) * kLanes];
      }
      const U4 rc = draw(seed, m, chain, segment, PMMC_LEVEL_HEATMAP, kSubChain);
      const float dx = disp(rc.x, step), dy = disp(rc.y, step);
      const float dz = S.use2d ? 0.f : disp(rc.z, step);
      const float nx = move_coord(px, dx), ny = move_coord(py, dy), nz = move_coord(pz, dz);
      float dE = hreg_delta<NB, ASYM>(lw, lw2, X, Y, Z, p, px, py, pz, nx, ny, nz);
      if (r > 0.f) dE += sphere_penalty(nx, ny, nz, r) - sphere_penalty(px, py, pz, r);
      const double En = E + static_cast<double>(dE);
      const bool ok = active && metropolis(S, E, En, T, u01(rc.w));
      if (ok) {
        g[(3 * p + 0) * kLanes] = nx;
        g[(3 * p + 1) * kLanes] = ny;
        g[(3 * p + 2) * kLanes] = nz;
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const bool hit = ok && j == p;
        X[j] = hit ? nx : X[j];
        Y[j] = hit ? ny : Y[j];
        Z[j] = hit ? nz : Z[j];
      }
      // the next move's prefetched copy was read before this move's store
      const bool same = ok && pn == p;
      qx = same ? nx : qx;
      qy = same ? ny : qy;
      qz = same ? nz : qz;
      if (ok) {
        E = En;
        ++msucc;
        ++acc;
        best = E < best ? E : best;
      }
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
      if (__ballot(due && resum_now) != 0) {
        Er = hreg_resum(P, it.pos, lane, r);
      }
      if (due) {
        ++miles;
        E = Er;
        milestone(S, E, ms, msucc, misses, status, active);
      }
    }
  }
  // the memory copy of the positions is already current
  if (status == -1) return;  // padding lane of a partial ensemble
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

// Lane-pair form for symmetric tables: a chain is held by two adjacent lanes,
// the even lane keeping beads 0, 2, .., 62 and the odd lane beads 1, 3, ..,
// 63 (96 position VGPRs instead of 192, so 3 waves per SIMD instead of 2).
// The move's ΔE splits exactly along that line: the even partners' terms go
// to acc0 and the odd partners' to acc1, each in increasing partner order, as
// in hreg_delta, so each lane computes one accumulator with the same
// roundings and one swap gives both lanes acc0 + acc1 (IEEE addition is
// commutative).  Everything else -- random streams, the Metropolis decision,
// the schedule -- is computed identically in both lanes of a pair.  One
// workgroup of two waves runs one 64-chain ensemble (wave w: chains 32w ..
// 32w + 31) and shares one LDS copy of W, rows de-interleaved: row p holds
// its even columns at [0, 32) and its odd columns at [kPairOdd, kPairOdd +
// 32), so the two lanes of a pair read different LDS banks.
constexpr int kPairOdd = 36;
constexpr int kPairRow = kPairOdd + 32;

__device__ __forceinline__ float hpair_delta(const float *lrow, const float (&X)[32],
                                             const float (&Y)[32], const float (&Z)[32],
                                             float px, float py, float pz, float nx, float ny,
                                             float nz) {
  const float4 *w1 = reinterpret_cast<const float4 *>(lrow);
  float acc = 0.f;
#pragma unroll
  for (int q = 0; q < 32; q += 4) {
    const float4 w = w1[q / 4];
    const float wq[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int a = q + u;
      const float a0 = fmaf(dist3(px, py, pz, X[a], Y[a], Z[a]), wq[u], -1.f);
      const float a1 = fmaf(dist3(nx, ny, nz, X[a], Y[a], Z[a]), wq[u], -1.f);
      acc = fmaf(a1 - a0, a1 + a0, acc);
    }
  }
  // acc0 + acc1 in both lanes of the pair: DPP quad_perm [1,0,3,2] swaps
  // adjacent lanes in one VALU op (no LDS round trip)
  const float other =
      __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(acc), 0xB1, 0xF, 0xF, false));
  return 2.f * (acc + other);
}

__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_anneal_hpair(
    const DevProblem *__restrict__ probs, const DevEns *__restrict__ ens, int max_blocks) {
  extern __shared__ float lw[];  // 64 rows of kPairRow floats
  const DevEns it = ens[blockIdx.x];
  const DevProblem &P = probs[it.problem];
  const int n = P.n;
  // the weight table first: no wave leaves before the barrier.  W rows are
  // 64 floats (P.ldw == 64: hreg_halves admits only such problems), zero
  // past column n; rows n..63 of the LDS table are zero.
  for (int i = threadIdx.x; i < 64 * kPairRow; i += 128) {
    const int row = i / kPairRow, c = i % kPairRow;
    const int col = c < 32 ? 2 * c : (c >= kPairOdd ? 2 * (c - kPairOdd) + 1 : -1);
    lw[i] = (row < n && col >= 0) ? P.W[row * 64 + col] : 0.f;
  }
  __syncthreads();

  const int half = threadIdx.x & 1;
  const int cl = threadIdx.x >> 1;  // chain within the ensemble, 0..63
  EnsState *__restrict__ st = it.st;
  int status = st->status[cl];
  if (__ballot(status == 0) == 0) return;

  float PMMC_AS_GLOBAL *g = (float PMMC_AS_GLOBAL *)(it.pos) + cl;
  float X[32], Y[32], Z[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const int j = 2 * i + half;
    X[i] = j < n ? g[(3 * j + 0) * kLanes] : 0.f;
    Y[i] = j < n ? g[(3 * j + 1) * kLanes] : 0.f;
    Z[i] = j < n ? g[(3 * j + 2) * kLanes] : 0.f;
  }

  const pmmc_schedule S = P.sched;
  const uint64_t seed = P.seed;
  const uint32_t segment = P.segment;
  const int nfree = P.nfree;
  const float r = S.sphere_radius;
  const float *lhalf = lw + (half ? kPairOdd : 0);

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
    U4 rbn = draw(seed, m0 >> 2, group, segment, PMMC_LEVEL_HEATMAP, kSubBead);
    int pn = cst(P.free_list)[pick(pick_word(rbn, m0), static_cast<uint32_t>(nfree))];
    float qx = g[(3 * pn + 0) * kLanes], qy = g[(3 * pn + 1) * kLanes], qz = g[(3 * pn + 2) * kLanes];
    for (int k = 0; k < S.block; ++k) {
      const uint32_t m = m0 + static_cast<uint32_t>(k);
      const int p = pn;
      const float px = qx, py = qy, pz = qz;
      const bool more = k + 1 < S.block;
      if (more) {
        const uint32_t mn = m + 1;
        if ((mn & 3u) == 0) rbn = draw(seed, mn >> 2, group, segment, PMMC_LEVEL_HEATMAP, kSubBead);
        pn = cst(P.free_list)[pick(pick_word(rbn, mn), static_cast<uint32_t>(nfree))];
        qx = g[(3 * pn + 0) * kLanes];
        qy = g[(3 * pn + 1) * kLanes];
        qz = g[(3 * pn + 2) * kLanes];
      }
      const U4 rc = draw(seed, m, chain, segment, PMMC_LEVEL_HEATMAP, kSubChain);
      const float dx = disp(rc.x, step), dy = disp(rc.y, step);
      const float dz = S.use2d ? 0.f : disp(rc.z, step);
      const float nx = move_coord(px, dx), ny = move_coord(py, dy), nz = move_coord(pz, dz);
      float dE = hpair_delta(lhalf + p * kPairRow, X, Y, Z, px, py, pz, nx, ny, nz);
      if (r > 0.f) dE += sphere_penalty(nx, ny, nz, r) - sphere_penalty(px, py, pz, r);
      const double En = E + static_cast<double>(dE);
      const bool ok = active && metropolis(S, E, En, T, u01(rc.w));
      if (ok && half == 0) {
        g[(3 * p + 0) * kLanes] = nx;
        g[(3 * p + 1) * kLanes] = ny;
        g[(3 * p + 2) * kLanes] = nz;
      }
      const bool mine = ok && (p & 1) == half;
      const int pi = p >> 1;
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        const bool hit = mine && i == pi;
        X[i] = hit ? nx : X[i];
        Y[i] = hit ? ny : Y[i];
        Z[i] = hit ? nz : Z[i];
      }
      const bool same = ok && pn == p;
      qx = same ? nx : qx;
      qy = same ? ny : qy;
      qz = same ? nz : qz;
      if (ok) {
        E = En;
        ++msucc;
        ++acc;
        best = E < best ? E : best;
      }
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
      if (__ballot(due && resum_now) != 0) {
        Er = hreg_resum(P, it.pos, cl, r);
      }
      if (due) {
        ++miles;
        E = Er;
        milestone(S, E, ms, msucc, misses, status, active);
      }
    }
  }
  if (status == -1 || half != 0) return;  // padding lane, or the pair's second lane
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

// heatmaps of 33..64 beads (weight rows padded to 64).  Up to 32 beads the
// partner loop over L2-resident positions is faster (N=24: 16.3 vs 10.3
// G moves/s).
int hreg_halves(const HostProblem &h) {
  if (h.level != PMMC_LEVEL_HEATMAP || h.n <= 32 || h.n > 64) return 0;
  return h.ldw == 64 ? 2 : 0;
}

// ensembles resident per CU: one 64-lane workgroup per ensemble (asymmetric
// tables), one 128-lane lane-pair workgroup per ensemble (symmetric)
int hreg_resident_per_cu(bool asym) {
  int blocks = 0;
  const hipError_t e =
      asym ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_anneal_hreg<64, true>, kLanes,
                                                          static_cast<size_t>(64) * 64 * sizeof(float) * 2)
           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_anneal_hpair, 2 * kLanes,
                                                          static_cast<size_t>(64) * kPairRow * sizeof(float));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return blocks;
}

template <int NB, bool ASYM>
static void launch_hreg(hipStream_t s, const DevProblem *probs, const DevEns *ens, int n_ens,
                        int max_blocks) {
  const size_t lds = static_cast<size_t>(NB) * NB * sizeof(float) * (ASYM ? 2 : 1);
  hipLaunchKernelGGL((k_anneal_hreg<NB, ASYM>), dim3(n_ens), dim3(kLanes), lds, s, probs, ens,
                     max_blocks);
}

hipError_t launch_heatmap_reg(hipStream_t s, const DevProblem *probs, const DevEns *ens, int n_ens,
                              int max_blocks, int halves, bool asym) {
  if (n_ens <= 0) return hipSuccess;
  if (halves != 2) return hipErrorInvalidValue;
  if (asym)
    launch_hreg<64, true>(s, probs, ens, n_ens, max_blocks);
  else
    hipLaunchKernelGGL(k_anneal_hpair, dim3(n_ens), dim3(2 * kLanes),
                       static_cast<size_t>(64) * kPairRow * sizeof(float), s, probs, ens, max_blocks);
  return hipGetLastError();
}

}  // namespace pmmc
