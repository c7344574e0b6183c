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
//
// Chains draw their bead picks per chain, so the lanes of a wavefront are
// independent, and a slot (a lane, or a lane pair) whose chain has stopped
// takes the next chain of its problem from a queue: with more chains than
// slots a wavefront no longer idles until its longest chain ends.  The
// workgroups of the problem's first P.workers ensembles are the workers;
// the others exit at once.  Chain c keeps its state in ensemble c/64, lane
// c%64 (the same place as without the queue), so results depend only on the
// chains, never on which slot ran them.
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
// A chain's positions seen through a per-lane pointer to its column of the
// SoA memory copy: element (bead j, axis k) at q[(3 j + k) * 64].
struct LaneView {
  const float PMMC_AS_GLOBAL *q;
  __device__ __forceinline__ float get(int j, int k) const { return q[(3 * j + k) * kLanes]; }
};

__device__ __noinline__ double hreg_resum(const DevProblem &P, const float PMMC_AS_GLOBAL *q, float r) {
  const LaneView V{q};
  return resum_heatmap(P, V, r);
}

// Bead of a pick index: the free-bead list, or the index itself when no bead
// is fixed (a uniform branch; the list is then the identity)
__device__ __forceinline__ int free_bead(const DevProblem &P, bool all_free, uint32_t i) {
  if (all_free) return static_cast<int>(i);
  return P.free_list[i];
}

// Per-chain annealing state (one EnsState lane) held in registers.
struct ChainRegs {
  double E, T, ms, best;
  float step;
  uint32_t iters, acc;
  int msucc, misses, miles, status;
};

__device__ __forceinline__ void chain_load(const EnsState *st, int l, ChainRegs &r) {
  r.E = st->E[l];
  r.T = st->T[l];
  r.ms = st->ms[l];
  r.best = st->best[l];
  r.step = st->step[l];
  r.iters = st->iters[l];
  r.acc = st->acc[l];
  r.msucc = st->msucc[l];
  r.misses = st->misses[l];
  r.miles = st->milestones[l];
  r.status = st->status[l];
}

__device__ __forceinline__ void chain_save(EnsState *st, int l, const ChainRegs &r) {
  st->E[l] = r.E;
  st->T[l] = r.T;
  st->ms[l] = r.ms;
  st->best[l] = r.best;
  st->step[l] = r.step;
  st->iters[l] = r.iters;
  st->acc[l] = r.acc;
  st->msucc[l] = r.msucc;
  st->misses[l] = r.misses;
  st->milestones[l] = r.miles;
  st->status[l] = r.status;
}

// Occupancy target: 2 waves per SIMD (192 position VGPRs of 256).
template <int NB, bool ASYM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_anneal_hreg(const DevProblem *__restrict__ probs,
                                                    const DevEns *__restrict__ ens, int max_blocks) {
  extern __shared__ float lw[];  // W (and W2), NB*NB floats each
  const DevEns it = ens[blockIdx.x];
  const DevProblem &P = probs[it.problem];
  const int w = static_cast<int>((it.chain0 - P.chain0) >> 6);  // worker index
  if (w >= P.workers) return;  // not a worker: its chains are run by the queue
  const int n = P.n;
  constexpr int L = NB;  // == P.ldw (rows padded with zero weights)
  float *lw2 = lw + L * L;
  // W and W2 hold n rows of L floats; rows n..L-1 of the LDS tables are zero
  for (int i = threadIdx.x; i < L * L; i += kLanes) {
    const bool in = i < n * L;
    lw[i] = in ? P.W[i] : 0.f;
    if constexpr (ASYM) lw2[i] = in ? P.W2[i] : 0.f;
  }
  __syncthreads();

  const DevEns *pe = ens + (static_cast<int>(blockIdx.x) - w);  // the problem's ensembles
  const int lane = threadIdx.x;
  const int slot = w * kLanes + lane;
  const pmmc_schedule S = P.sched;
  const uint64_t seed = P.seed;
  const uint32_t segment = P.segment;
  const int nfree = P.nfree;
  const float r = S.sphere_radius;
  const int nch = P.n_chains;
  const bool all_free = nfree == n;
  const bool aligned4 = (S.block & 3) == 0;

  // the slot's chain: memory copy of its positions (global address space,
  // not generic: flat loads and stores would also count on lgkmcnt, so every
  // wait for a weight-row LDS read would wait for the copy's traffic too) and
  // its state; lanes without a chain keep a valid address (their own
  // ensemble slot) and never store
  int c = P.slot_chain[slot];
  bool drained = c == -2, have = false, active = false;
  float PMMC_AS_GLOBAL *g = (float PMMC_AS_GLOBAL *)(it.pos) + lane;
  ChainRegs R{};
  float X[NB], Y[NB], Z[NB];

  for (int b = 0; b < max_blocks; ++b) {
    // Every slot ends this loop running a chain, or with the queue drained:
    // a stopped chain's state goes back and the slot takes the next chain.
    for (;;) {
      if (have && active && static_cast<int64_t>(R.iters) >= S.max_moves) {
        R.status = PMMC_STOP_CAP;
        active = false;
      }
      if (have && !active) {
        chain_save(pe[c >> 6].st, c & (kLanes - 1), R);
        have = false;
        c = -1;
      }
      if (have || drained) break;
      if (c < 0) {
        const int t = atomicAdd(P.queue, 1);
        if (t >= nch) {
          drained = true;
          break;
        }
        c = t;
      }
      // start (or resume) chain c
      const DevEns ce = pe[c >> 6];
      g = (float PMMC_AS_GLOBAL *)(ce.pos) + (c & (kLanes - 1));
      chain_load(ce.st, c & (kLanes - 1), R);
      have = true;
      active = R.status == 0;
    }
    if (__ballot(active) == 0) break;
    // the registers take the chain's positions from its memory copy (always
    // current) at every block: no merge of old and new registers on a refill
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      X[j] = j < n ? g[(3 * j + 0) * kLanes] : 0.f;
      Y[j] = j < n ? g[(3 * j + 1) * kLanes] : 0.f;
      Z[j] = j < n ? g[(3 * j + 2) * kLanes] : 0.f;
    }
    const uint32_t m0 = R.iters;
    // bead picks per chain (pick group = chain: the host sets pick_shift 0
    // for every problem these kernels run)
    const uint32_t chain = P.chain0 + static_cast<uint32_t>(c < 0 ? 0 : c);
    const uint32_t group = chain;
    // The moved bead's old position comes from the SoA copy in memory (kept
    // current on every accepted move), read one move ahead: the pick stream
    // does not depend on the chains' state.
    U4 rbn = draw(seed, m0 >> 2, group, segment, PMMC_LEVEL_HEATMAP, kSubBead);
    int pn = free_bead(P, all_free, pick(pick_word(rbn, m0), static_cast<uint32_t>(nfree)));
    float qx = g[(3 * pn + 0) * kLanes], qy = g[(3 * pn + 1) * kLanes], qz = g[(3 * pn + 2) * kLanes];
    double E = R.E, T = R.T;
    const float step = R.step;
    for (int k = 0; k < S.block; ++k) {
      const uint32_t m = m0 + static_cast<uint32_t>(k);
      const int p = pn;
      const float px = qx, py = qy, pz = qz;
      const bool more = k + 1 < S.block;
      if (more) {
        const uint32_t mn = m + 1;
        // a chain's move count is a multiple of the block, so with blocks of a
        // multiple of 4 moves the redraw is the same in every lane
        if (aligned4 ? ((k + 1) & 3) == 0 : (mn & 3u) == 0)
          rbn = draw(seed, mn >> 2, group, segment, PMMC_LEVEL_HEATMAP, kSubBead);
        pn = free_bead(P, all_free, pick(pick_word(rbn, mn), static_cast<uint32_t>(nfree)));
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
      if (__ballot(ok) != 0) {  // wave-uniform: skipped when every chain rejects
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          const bool hit = ok && j == p;
          X[j] = hit ? nx : X[j];
          Y[j] = hit ? ny : Y[j];
          Z[j] = hit ? nz : Z[j];
        }
      }
      // the next move's prefetched copy was read before this move's store
      const bool same = ok && pn == p;
      qx = same ? nx : qx;
      qy = same ? ny : qy;
      qz = same ? nz : qz;
      if (ok) {
        E = En;
        ++R.msucc;
        ++R.acc;
        R.best = E < R.best ? E : R.best;
      }
      if (active && !S.cool_per_block) T *= S.dt;
    }
    R.E = E;
    R.T = T;
    bool due = false;
    if (active) {
      R.iters += static_cast<uint32_t>(S.block);
      if (S.cool_per_block) R.T *= S.dt;
      R.step *= S.step_decay;
      due = (R.iters % static_cast<uint32_t>(S.milestone_period)) < static_cast<uint32_t>(S.block);
    }
    if (__ballot(due) != 0) {
      double Er = R.E;
      const bool resum_now = S.resum > 0 && ((R.miles + 1) % S.resum) == 0;
      if (__ballot(due && resum_now) != 0) {
        Er = hreg_resum(P, g, r);
      }
      if (due) {
        ++R.miles;
        R.E = Er;
        milestone(S, R.E, R.ms, R.msucc, R.misses, R.status, active);
      }
    }
  }
  // the memory copy of the positions is already current
  if (have) chain_save(pe[c >> 6].st, c & (kLanes - 1), R);
  P.slot_chain[slot] = have ? c : (drained ? -2 : -1);
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
  const int w = static_cast<int>((it.chain0 - P.chain0) >> 6);  // worker index
  if (w >= P.workers) return;  // not a worker: its chains are run by the queue
  const int n = P.n;
  // the weight table first: no wave leaves before the barrier.  W rows are
  // 64 floats (P.ldw == 64: hreg_halves admits only such problems), zero
  // past column n; rows n..63 of the LDS table are zero.
  for (int i = threadIdx.x; i < 64 * kPairRow; i += 128) {
    const int row = i / kPairRow, cc = i % kPairRow;
    const int col = cc < 32 ? 2 * cc : (cc >= kPairOdd ? 2 * (cc - kPairOdd) + 1 : -1);
    lw[i] = (row < n && col >= 0) ? P.W[row * 64 + col] : 0.f;
  }
  __syncthreads();

  const DevEns *pe = ens + (static_cast<int>(blockIdx.x) - w);  // the problem's ensembles
  const int half = threadIdx.x & 1;
  const int cl = threadIdx.x >> 1;  // slot within the worker, 0..63
  const int wl = threadIdx.x & (kLanes - 1);  // lane within the wavefront
  const int slot = w * kLanes + cl;
  const pmmc_schedule S = P.sched;
  const uint64_t seed = P.seed;
  const uint32_t segment = P.segment;
  const int nfree = P.nfree;
  const float r = S.sphere_radius;
  const int nch = P.n_chains;
  const bool all_free = nfree == n;
  const bool aligned4 = (S.block & 3) == 0;
  const float *lhalf = lw + (half ? kPairOdd : 0);

  // The slot's chain (both lanes of a pair run the same one): memory copy of
  // its positions (global address space) and its state.  Lanes without a
  // chain keep a valid address (their own ensemble slot) and never store.
  int c = P.slot_chain[slot];
  bool drained = c == -2, have = false, active = false;
  float PMMC_AS_GLOBAL *g = (float PMMC_AS_GLOBAL *)(it.pos) + cl;
  ChainRegs R{};
  float X[32], Y[32], Z[32];

  for (int b = 0; b < max_blocks; ++b) {
    // Every slot ends this loop running a chain, or with the queue drained;
    // the two lanes of a pair take the same path.
    for (;;) {
      if (have && active && static_cast<int64_t>(R.iters) >= S.max_moves) {
        R.status = PMMC_STOP_CAP;
        active = false;
      }
      if (have && !active) {
        if (half == 0) chain_save(pe[c >> 6].st, c & (kLanes - 1), R);
        have = false;
        c = -1;
      }
      if (have || drained) break;
      if (c < 0) {
        int t = 0;
        if (half == 0) t = atomicAdd(P.queue, 1);
        t = __shfl(t, wl & ~1, kLanes);
        if (t >= nch) {
          drained = true;
          break;
        }
        c = t;
      }
      const DevEns ce = pe[c >> 6];
      g = (float PMMC_AS_GLOBAL *)(ce.pos) + (c & (kLanes - 1));
      chain_load(ce.st, c & (kLanes - 1), R);
      have = true;
      active = R.status == 0;
    }
    if (__ballot(active) == 0) break;
    // registers from the chain's memory copy (always current), every block
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      const int j = 2 * i + half;
      X[i] = j < n ? g[(3 * j + 0) * kLanes] : 0.f;
      Y[i] = j < n ? g[(3 * j + 1) * kLanes] : 0.f;
      Z[i] = j < n ? g[(3 * j + 2) * kLanes] : 0.f;
    }
    const uint32_t m0 = R.iters;
    // bead picks per chain (pick group = chain: the host sets pick_shift 0
    // for every problem these kernels run)
    const uint32_t chain = P.chain0 + static_cast<uint32_t>(c < 0 ? 0 : c);
    const uint32_t group = chain;
    U4 rbn = draw(seed, m0 >> 2, group, segment, PMMC_LEVEL_HEATMAP, kSubBead);
    int pn = free_bead(P, all_free, pick(pick_word(rbn, m0), static_cast<uint32_t>(nfree)));
    float qx = g[(3 * pn + 0) * kLanes], qy = g[(3 * pn + 1) * kLanes], qz = g[(3 * pn + 2) * kLanes];
    double E = R.E, T = R.T;
    const float step = R.step;
    for (int k = 0; k < S.block; ++k) {
      const uint32_t m = m0 + static_cast<uint32_t>(k);
      const int p = pn;
      const float px = qx, py = qy, pz = qz;
      const bool more = k + 1 < S.block;
      if (more) {
        const uint32_t mn = m + 1;
        // a chain's move count is a multiple of the block, so with blocks of a
        // multiple of 4 moves the redraw is the same in every lane
        if (aligned4 ? ((k + 1) & 3) == 0 : (mn & 3u) == 0)
          rbn = draw(seed, mn >> 2, group, segment, PMMC_LEVEL_HEATMAP, kSubBead);
        pn = free_bead(P, all_free, pick(pick_word(rbn, mn), static_cast<uint32_t>(nfree)));
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
      // a wave-uniform branch: most moves are rejected in every chain of the
      // wave, and then the 32 x (compare + 3 selects) are skipped
      if (__ballot(mine) != 0) {
#pragma unroll
        for (int i = 0; i < 32; ++i) {
          const bool hit = mine && i == pi;
          X[i] = hit ? nx : X[i];
          Y[i] = hit ? ny : Y[i];
          Z[i] = hit ? nz : Z[i];
        }
      }
      const bool same = ok && pn == p;
      qx = same ? nx : qx;
      qy = same ? ny : qy;
      qz = same ? nz : qz;
      if (ok) {
        E = En;
        ++R.msucc;
        ++R.acc;
        R.best = E < R.best ? E : R.best;
      }
      if (active && !S.cool_per_block) T *= S.dt;
    }
    R.E = E;
    R.T = T;
    bool due = false;
    if (active) {
      R.iters += static_cast<uint32_t>(S.block);
      if (S.cool_per_block) R.T *= S.dt;
      R.step *= S.step_decay;
      due = (R.iters % static_cast<uint32_t>(S.milestone_period)) < static_cast<uint32_t>(S.block);
    }
    if (__ballot(due) != 0) {
      double Er = R.E;
      const bool resum_now = S.resum > 0 && ((R.miles + 1) % S.resum) == 0;
      if (__ballot(due && resum_now) != 0) {
        Er = hreg_resum(P, g, r);
      }
      if (due) {
        ++R.miles;
        R.E = Er;
        milestone(S, R.E, R.ms, R.msucc, R.misses, R.status, active);
      }
    }
  }
  if (half != 0) return;  // the pair's second lane
  if (have) chain_save(pe[c >> 6].st, c & (kLanes - 1), R);
  P.slot_chain[slot] = have ? c : (drained ? -2 : -1);
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
