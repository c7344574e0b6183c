// kernels.hpp -- device code of the lockstep-ensemble annealer (gfx950).
//
// One workgroup = one wavefront = one ensemble of 64 chains (lane = chain).
// Per move, the bead index is wave-uniform (shared pick stream), so all
// static problem data for the move is read through scalar loads, while each
// lane reads its own copy of the partner positions from a coalesced SoA row.
// There is no cross-lane reduction on the move path: every lane sums its own
// chain's partners; ballots only decide when the ensemble has finished.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "device.hpp"
#include "engine.hpp"

namespace pmmc {

#ifndef PMMC_ARCS_CHUNK
#define PMMC_ARCS_CHUNK 8
#endif
// arc CSR entries evaluated per round trip of the move's partner loop
constexpr int kArcsChunk = PMMC_ARCS_CHUNK;

#ifndef PMMC_ARCS_DEPTH
#define PMMC_ARCS_DEPTH 1
#endif
// arc-level moves whose rows and positions are in flight ahead of the one
// evaluated.  Depth 2 (140 VGPRs, 3 waves per SIMD) measured slower on every
// form: 400-anchor solve 5.91 vs 5.98e10 moves/s, 512-problem pool 2,639 vs
// 2,840 solves/s, lone plain solves 12.9 vs 14.5 per second.
constexpr int kArcsDepth = PMMC_ARCS_DEPTH;
static_assert(kArcsDepth >= 1, "at least the next move is fetched ahead");

#ifndef PMMC_SMOOTH_DEPTH
#define PMMC_SMOOTH_DEPTH 1
#endif
// smooth-level moves whose windows are in flight ahead of the one evaluated.
// Measured on the 19,951-bead smooth segment: 74.0 / 71.7 / 71.0 G moves/s
// at depth 1 / 2 / 3 (each window costs 17 VGPRs, and depth 2 drops the
// kernel from 5 to 4 waves per SIMD), so one move ahead is the default.
constexpr int kSmoothDepth = PMMC_SMOOTH_DEPTH;
static_assert(kSmoothDepth >= 1, "at least the next move is fetched ahead");

// ------------------------------------------------------------------ //
// helpers
// ------------------------------------------------------------------ //
__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int lane) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), lane));
}

#define PMMC_AS_GLOBAL __attribute__((address_space(1)))
#define PMMC_AS_CONST __attribute__((address_space(4)))

// A problem's static tables (free list, CSR, weight rows, dist_to_next) are
// never written while a kernel runs.  Reading them through the constant
// address space lets the compiler use scalar loads for wave-uniform
// addresses (the moved bead and its CSR row are wave-uniform).
template <class T>
__device__ __forceinline__ const PMMC_AS_CONST T *cst(const T *p) {
  return (const PMMC_AS_CONST T *)p;
}

// One CSR entry / pair entry read field by field through the constant
// address space (scalar loads for a wave-uniform index).
__device__ __forceinline__ ArcEntry load_entry(const ArcEntry *t, int i) {
  const PMMC_AS_CONST ArcEntry *q = cst(t) + i;
  ArcEntry e;
  e.j = q->j;
  e.c = q->c;
  e.ka = q->ka;
  e.kb = q->kb;
  return e;
}

__device__ __forceinline__ PairEntry load_pair(const PairEntry *t, int i) {
  const PMMC_AS_CONST PairEntry *q = cst(t) + i;
  PairEntry e;
  e.i = q->i;
  e.e.j = q->e.j;
  e.e.c = q->e.c;
  e.e.ka = q->e.ka;
  e.e.kb = q->e.kb;
  return e;
}

// Position accessors: element (bead j, axis k) of this lane's chain.
template <bool LDS>
struct PosViewT;

// positions in HBM: the ensemble's base (wave-uniform) and this lane's column,
// so each partner costs one address computation and its three axes are
// immediate offsets of one global load form
template <>
struct PosViewT<false> {
  float PMMC_AS_GLOBAL *base;
  unsigned lane;
  __device__ __forceinline__ float get(int j, int k) const {
    return (base + 3 * j * kLanes)[k * kLanes + lane];
  }
  // write-backs are non-temporal stores: a written row is rarely read again
  // before the L2 evicts it (arc level, 400 anchors: 60.83 -> 61.32 G
  // moves/s, three A/B pairs on one GPU)
  __device__ __forceinline__ void set(int j, int k, float v) const {
    __builtin_nontemporal_store(v, &(base + 3 * j * kLanes)[k * kLanes + lane]);
  }
};

// positions staged in LDS: p = lds + lane
template <>
struct PosViewT<true> {
  float *p;
  __device__ __forceinline__ float get(int j, int k) const { return p[(3 * j + k) * kLanes]; }
  __device__ __forceinline__ void set(int j, int k, float v) const { p[(3 * j + k) * kLanes] = v; }
};

using PosView = PosViewT<false>;

__device__ __forceinline__ PosView global_view(float *pos, unsigned lane) {
  PosView v;
  v.base = (float PMMC_AS_GLOBAL *)pos;
  v.lane = lane;
  return v;
}

// An ensemble's positions in HBM are laid out per level.  Heatmap and arc
// levels: SoA rows, element (bead j, axis k) of lane l at (3j + k) * 64 + l,
// so each axis of a bead is one 256-byte row.  Smooth level: a lane's three
// coordinates together, at (64j + l) * 3 + k; a bead is still one 768-byte
// span, read by one 12-byte load per lane, and a move accepted by one lane
// (the smooth level's 1.4% acceptance) dirties one cache line instead of
// three: the move pattern with one lane writing in about every other move
// sustains 89.7 instead of 81.0 G moves/s.  With about half of the lanes
// writing (the arc level's 51%) the SoA rows measured 1.6% faster.
template <int LEVEL>
constexpr bool kLaneXyz = LEVEL == PMMC_LEVEL_SMOOTH;

__host__ __device__ __forceinline__ size_t pos_index(bool lane_xyz, int j, int k, unsigned lane) {
  return lane_xyz ? (static_cast<size_t>(j) * kLanes + lane) * 3 + k
                  : static_cast<size_t>(3 * j + k) * kLanes + lane;
}

// positions in HBM, lane-contiguous coordinates: base = pos + 3 * lane
struct PosViewXyz {
  float PMMC_AS_GLOBAL *base;
  __device__ __forceinline__ float get(int j, int k) const { return base[j * (3 * kLanes) + k]; }
  __device__ __forceinline__ void set(int j, int k, float v) const { base[j * (3 * kLanes) + k] = v; }
};

template <int LEVEL>
using GlobalViewT = typename std::conditional<kLaneXyz<LEVEL>, PosViewXyz, PosView>::type;

template <int LEVEL>
__device__ __forceinline__ GlobalViewT<LEVEL> level_view(float *pos, unsigned lane) {
  if constexpr (kLaneXyz<LEVEL>) {
    PosViewXyz v;
    v.base = (float PMMC_AS_GLOBAL *)pos + 3 * lane;
    return v;
  } else {
    return global_view(pos, lane);
  }
}

// Pick word of move m from the draw of its group of four moves.
__device__ __forceinline__ uint32_t pick_word(const U4 &rb, uint32_t m) {
  const uint32_t sel = m & 3u;
  return sel == 0 ? rb.x : (sel == 1 ? rb.y : (sel == 2 ? rb.z : rb.w));
}

__device__ __forceinline__ bool metropolis(const pmmc_schedule &S, double E, double En, double T,
                                           float U) {
  bool ok = S.strict_accept ? (En < E) : (En <= E);
  if (!ok && T > 0.0 && (!S.positive_guard || E > 0.0)) {
    // tp = tjs * exp(-tjc * (E'/E) / T)   (ParallelMonteCarloArcs.cu:600-603)
    const float ratio = static_cast<float>(En / E);
    const float tp = S.tjs * __expf(-S.tjc * ratio / static_cast<float>(T));
    ok = U < tp;
  }
  return ok;
}

// Milestone stop test on the (re-summed) running total E: the
// no-improvement predicate and the three stop rules of the reference
// annealers (LooperSolver.hpp:1543-1553, :4765-4783, :5015-5034;
// ParallelMonteCarloHeatmap.cu:334-352; ParallelMonteCarloArcs.cu:636-655;
// ParallelMonteCarloSmooth.cu:320-335).  Updates ms/msucc when the chain
// continues.
__device__ __forceinline__ void milestone(const pmmc_schedule &S, double E, double &ms, int &msucc,
                                          int &misses, int &status, bool &active) {
  if (!isfinite(E)) {
    status = PMMC_STOP_NONFINITE;
    active = false;
    return;
  }
  const bool no_imp = (E > S.improvement * ms && msucc < S.min_successes) || E < S.small_energy ||
                      (S.ratio_stop > 0.0 && ms > 1e-10 && E / ms > S.ratio_stop);
  if (S.stop_rule == PMMC_STOP_FIRST_FAIL) {
    if (no_imp) {
      status = E < S.small_energy ? PMMC_STOP_SMALL : PMMC_STOP_RULE;
      active = false;
    }
  } else {
    if (no_imp) ++misses;
    else if (S.stop_rule == PMMC_STOP_MISSES_RESET) misses = 0;
    if (E < S.small_energy) {
      status = PMMC_STOP_SMALL;
      active = false;
    } else if (misses >= S.fails_threshold) {
      status = PMMC_STOP_RULE;
      active = false;
    }
  }
  if (active) {
    ms = E;
    msucc = 0;
  }
}

// ------------------------------------------------------------------ //
// level energies, FP32 move deltas
// ------------------------------------------------------------------ //

#ifndef PMMC_HEAT_CHUNK
#define PMMC_HEAT_CHUNK 16
#endif
// partner positions loaded per round trip of the heatmap partner loop
constexpr int kHeatChunk = PMMC_HEAT_CHUNK;
static_assert(kHeatChunk % 2 == 0 && kLanes % kHeatChunk == 0, "even chunk dividing 64");

// heatmap: exact change of the full (double-counted) score for moving p from
// xp to xn; LooperSolver.hpp:3583-3617 summed over all beads.
template <class View>
__device__ __forceinline__ float heatmap_delta(const DevProblem &P, const View &X, int p,
                                               float px, float py, float pz, float nx, float ny,
                                               float nz) {
  const int n = P.n;
  const int lane = threadIdx.x & (kLanes - 1);
  // The weight row of the (wave-uniform) bead p is read 64 entries at a time
  // with one coalesced vector load; partner j's weight is then broadcast from
  // lane j%64 with v_readlane.  Partners are taken kHeatChunk at a time: all
  // their position loads are issued before any is consumed.  Partner j's term
  // goes to accumulator j%2, in increasing j; past the last partner the chunk
  // is padded with weight 0 (a term of exactly 0).
  const float PMMC_AS_GLOBAL *w1 = (const float PMMC_AS_GLOBAL *)P.W + static_cast<size_t>(p) * P.ldw;
  float acc0 = 0.f, acc1 = 0.f;
  for (int c = 0; c < n; c += kLanes) {
    const int cnt = min(kLanes, n - c);
    const float wv = lane < cnt ? w1[c + lane] : 0.f;
    for (int jj = 0; jj < cnt; jj += kHeatChunk) {
      float xs[kHeatChunk], ys[kHeatChunk], zs[kHeatChunk];
#pragma unroll
      for (int u = 0; u < kHeatChunk; ++u) {
        const int j = min(c + jj + u, n - 1);
        xs[u] = X.get(j, 0);
        ys[u] = X.get(j, 1);
        zs[u] = X.get(j, 2);
      }
#pragma unroll
      for (int u = 0; u < kHeatChunk; u += 2) {
        const float wa = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wv), jj + u));
        const float wb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wv), jj + u + 1));
        const float a0 = fmaf(dist3(px, py, pz, xs[u], ys[u], zs[u]), wa, -1.f);
        const float a1 = fmaf(dist3(nx, ny, nz, xs[u], ys[u], zs[u]), wa, -1.f);
        const float b0 = fmaf(dist3(px, py, pz, xs[u + 1], ys[u + 1], zs[u + 1]), wb, -1.f);
        const float b1 = fmaf(dist3(nx, ny, nz, xs[u + 1], ys[u + 1], zs[u + 1]), wb, -1.f);
        acc0 = fmaf(a1 - a0, a1 + a0, acc0);
        acc1 = fmaf(b1 - b0, b1 + b0, acc1);
      }
    }
  }
  float d = acc0 + acc1;
  if (P.symmetric) return 2.f * d;
  const float PMMC_AS_GLOBAL *w2 = (const float PMMC_AS_GLOBAL *)P.W2 + static_cast<size_t>(p) * P.ldw;
  float acc2 = 0.f;
  for (int c = 0; c < n; c += kLanes) {
    const int cnt = min(kLanes, n - c);
    const float wv = lane < cnt ? w2[c + lane] : 0.f;
    for (int jj = 0; jj < cnt; jj += kHeatChunk) {
      float xs[kHeatChunk], ys[kHeatChunk], zs[kHeatChunk];
#pragma unroll
      for (int u = 0; u < kHeatChunk; ++u) {
        const int j = min(c + jj + u, n - 1);
        xs[u] = X.get(j, 0);
        ys[u] = X.get(j, 1);
        zs[u] = X.get(j, 2);
      }
#pragma unroll
      for (int u = 0; u < kHeatChunk; ++u) {
        const float w = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wv), jj + u));
        const float a0 = fmaf(dist3(px, py, pz, xs[u], ys[u], zs[u]), w, -1.f);
        const float a1 = fmaf(dist3(nx, ny, nz, xs[u], ys[u], zs[u]), w, -1.f);
        acc2 = fmaf(a1 - a0, a1 + a0, acc2);
      }
    }
  }
  return d + acc2;
}

// The arc-level ΔE in FP32 is defined as two accumulators: the terms of the
// even entries of bead p's CSR row (relative to its start) in increasing
// order, and those of the odd entries likewise, then acc[0] + acc[1].  A
// kernel that splits a chain's partners over two lanes by entry parity
// forms the same two sums, so every arc kernel gives the same bits.
// Term of one entry: (d1 - c)^2 k1 - (d0 - c)^2 k0, with d0 / d1 the
// distances from the old / new position of p to the partner, rounded as
// fma((d1 - c)^2, k1, -((d0 - c)^2 k0)) with no further contraction: the
// rounding is fixed in the source, not left to where the compiler fuses a
// product into a sum, which can differ between the kernels that inline it.
__device__ __forceinline__ float arcs_term(float c, float ka, float kb, float px, float py,
                                           float pz, float nx, float ny, float nz, float xj,
                                           float yj, float zj) {
#pragma clang fp contract(off)
  const float t0 = dist3(px, py, pz, xj, yj, zj) - c;
  const float t1 = dist3(nx, ny, nz, xj, yj, zj) - c;
  const float k0 = t0 >= 0.f ? ka : kb;
  const float k1 = t1 >= 0.f ? ka : kb;
  const float s0 = t0 * t0, s1 = t1 * t1;
  return fmaf(s1, k1, -(s0 * k0));
}

// One chunk (kArcsChunk entries, an even offset into the row) of bead p's
// entries: entry u adds to acc[u & 1], for the first `cnt` entries.
__device__ __forceinline__ void arcs_chunk(const ArcEntry (&en)[kArcsChunk], const float (&xj)[kArcsChunk],
                                           const float (&yj)[kArcsChunk], const float (&zj)[kArcsChunk],
                                           int cnt, float px, float py, float pz, float nx, float ny,
                                           float nz, float (&acc)[2]) {
  static_assert(kArcsChunk % 2 == 0, "chunks start at even entries");
#pragma unroll
  for (int u = 0; u < kArcsChunk; ++u) {
    if (u < cnt)
      acc[u & 1] += arcs_term(en[u].c, en[u].ka, en[u].kb, px, py, pz, nx, ny, nz, xj[u], yj[u], zj[u]);
  }
}

// Entries [e, e + kArcsChunk) of bead p (clamped to e1 - 1) and the positions
// of their partners.  The (wave-uniform) entry loads and then the
// 3*kArcsChunk per-lane position loads are all issued before any is consumed.
template <class View>
__device__ __forceinline__ void arcs_load_chunk(const DevProblem &P, const View &X, int e, int e1,
                                                ArcEntry (&en)[kArcsChunk], float (&xj)[kArcsChunk],
                                                float (&yj)[kArcsChunk], float (&zj)[kArcsChunk]) {
#pragma unroll
  for (int u = 0; u < kArcsChunk; ++u) en[u] = load_entry(P.ent, min(e + u, e1 - 1));
#pragma unroll
  for (int u = 0; u < kArcsChunk; ++u) {
    const int j = en[u].j & 0x7fffffff;
    xj[u] = X.get(j, 0);
    yj[u] = X.get(j, 1);
    zj[u] = X.get(j, 2);
  }
}

// arcs: change of arc + loop terms of bead p (ParallelMonteCarloArcs.cu:294-351,
// LooperSolver.hpp:3250-3281, :3355-3371).  Term (d-c)^2 * (d >= c ? ka : kb).
template <class View>
__device__ __forceinline__ float arcs_delta(const DevProblem &P, const View &X, int p, float px,
                                            float py, float pz, float nx, float ny, float nz) {
  const int e0 = cst(P.off)[p], e1 = cst(P.off)[p + 1];
  float acc[2] = {0.f, 0.f};
  for (int e = e0; e < e1; e += kArcsChunk) {
    ArcEntry en[kArcsChunk];
    float xj[kArcsChunk], yj[kArcsChunk], zj[kArcsChunk];
    arcs_load_chunk(P, X, e, e1, en, xj, yj, zj);
    arcs_chunk(en, xj, yj, zj, e1 - e, px, py, pz, nx, ny, nz, acc);
  }
  return acc[0] + acc[1];
}

// Everything one arc-level move reads from the chains' state before its
// decision, fetched ahead of time: the bead, its CSR range, the partner
// indices of its first kArcsChunk entries (slots past the range repeat the
// last entry) and the positions of the bead and of those partners.  The
// entries' coefficients are static and are read when the move is evaluated.
struct ArcsFetch {
  int p, e0, e1;
  int j[kArcsChunk];
  float px, py, pz;
  float xj[kArcsChunk], yj[kArcsChunk], zj[kArcsChunk];
};

// The bead picks of 64 consecutive moves mw .. mw + 63 and the first chunk of
// their CSR rows, gathered one move per lane (lane k holds move mw + k) with
// vector loads.  A move's fetch then reads them back with readlane, so the
// per-move chain of dependent scalar loads (free-list entry, row offsets,
// partner indices) becomes one gather per 64 moves.  Lanes past the block's
// last move gather picks nobody reads; every index stays in its table (the
// pick is below nfree, p below n, and the entry index is clamped to the row,
// or to entry 0 of a table holding at least one entry when the row is empty).
struct ArcsWindow {
  int p, e0, e1;
  int j[kArcsChunk];
};

__device__ __forceinline__ void arcs_window(const DevProblem &P, uint64_t seed, uint32_t group,
                                            uint32_t segment, uint32_t mw, ArcsWindow &w) {
  const uint32_t m = mw + static_cast<uint32_t>(threadIdx.x & (kLanes - 1));
  const U4 rb = draw(seed, m >> 2, group, segment, PMMC_LEVEL_ARCS, kSubBead);
  w.p = P.free_list[pick(pick_word(rb, m), static_cast<uint32_t>(P.nfree))];
  w.e0 = P.off[w.p];
  w.e1 = P.off[w.p + 1];
#pragma unroll
  for (int u = 0; u < kArcsChunk; ++u) w.j[u] = P.ej[max(min(w.e0 + u, w.e1 - 1), 0)] & 0x7fffffff;
}

// arcs_fetch of move mw + k from its window
template <class View>
__device__ __forceinline__ void arcs_fetch_win(const View &X, const ArcsWindow &w, int k,
                                               ArcsFetch &f) {
  f.p = __builtin_amdgcn_readlane(w.p, k);
  f.e0 = __builtin_amdgcn_readlane(w.e0, k);
  f.e1 = __builtin_amdgcn_readlane(w.e1, k);
  f.px = X.get(f.p, 0);
  f.py = X.get(f.p, 1);
  f.pz = X.get(f.p, 2);
  if (f.e1 > f.e0) {
#pragma unroll
    for (int u = 0; u < kArcsChunk; ++u) f.j[u] = __builtin_amdgcn_readlane(w.j[u], k);
#pragma unroll
    for (int u = 0; u < kArcsChunk; ++u) {
      f.xj[u] = X.get(f.j[u], 0);
      f.yj[u] = X.get(f.j[u], 1);
      f.zj[u] = X.get(f.j[u], 2);
    }
  }
}

// The prefetched move reads positions as they were before the current move;
// when the current move changed bead q (accepted per lane: ok), the copies of
// q among them take its new position.  Selects, no branches: a branch per
// slot makes every fetched value a merge and costs register copies.
__device__ __forceinline__ void arcs_patch(ArcsFetch &f, int q, bool ok, float qx, float qy,
                                           float qz) {
  const bool self = ok && f.p == q;
  f.px = self ? qx : f.px;
  f.py = self ? qy : f.py;
  f.pz = self ? qz : f.pz;
  const int cnt = min(kArcsChunk, f.e1 - f.e0);
#pragma unroll
  for (int u = 0; u < kArcsChunk; ++u) {
    const bool hit = ok && u < cnt && f.j[u] == q;
    f.xj[u] = hit ? qx : f.xj[u];
    f.yj[u] = hit ? qy : f.yj[u];
    f.zj[u] = hit ? qz : f.zj[u];
  }
}

// ΔE of a fetched move: the first chunk from the fetched positions, any
// further chunks (beads with more than kArcsChunk entries) read now; same
// terms and order as arcs_delta.  The first chunk's coefficients are read
// with scalar loads issued together before any term, entries past the row
// included (the tables carry kArcsChunk entries of padding): a load inside
// each entry's `u < cnt` branch made every entry wait for its own load.
// (Fetching them a move ahead instead costs 48 VGPRs, a wave per SIMD, and
// measured slower.)
template <class View>
__device__ __forceinline__ float arcs_delta_fetched(const DevProblem &P, const View &X,
                                                    const ArcsFetch &f, float nx, float ny,
                                                    float nz) {
  float acc[2] = {0.f, 0.f};
  const int cnt = min(kArcsChunk, f.e1 - f.e0);
  if (cnt > 0) {
    float c[kArcsChunk], ka[kArcsChunk], kb[kArcsChunk];
#pragma unroll
    for (int u = 0; u < kArcsChunk; ++u) {
      c[u] = cst(P.ec)[f.e0 + u];
      ka[u] = cst(P.eka)[f.e0 + u];
      kb[u] = cst(P.ekb)[f.e0 + u];
    }
#pragma unroll
    for (int u = 0; u < kArcsChunk; ++u) {
      if (u < cnt)
        acc[u & 1] += arcs_term(c[u], ka[u], kb[u], f.px, f.py, f.pz, nx, ny, nz, f.xj[u], f.yj[u],
                                f.zj[u]);
    }
  }
  for (int e = f.e0 + kArcsChunk; e < f.e1; e += kArcsChunk) {
    ArcEntry en[kArcsChunk];
    float xj[kArcsChunk], yj[kArcsChunk], zj[kArcsChunk];
    arcs_load_chunk(P, X, e, f.e1, en, xj, yj, zj);
    arcs_chunk(en, xj, yj, zj, f.e1 - e, f.px, f.py, f.pz, nx, ny, nz, acc);
  }
  return acc[0] + acc[1];
}

// smooth: bond term ((d - d0)/d0)^2 * k_{stretch|squeeze}, d0 = max(dtn, 1e-6)
__device__ __forceinline__ float bond_f(float d, float d0, float ks, float kq) {
  const float diff = (d - d0) / d0;
  return diff * diff * (diff >= 0.f ? ks : kq);
}

// angle term at a vertex with bond vectors u (prev -> vertex) and w (vertex ->
// next) given as u = a - b, w = b - c: ((1 - cos)/2)^3 * ka, skipped when a
// bond is shorter than 1e-6 (ParallelMonteCarloSmooth.cu:126-147).  1 - cos is
// evaluated as |u x w|^2 / (|u||w| (|u||w| + u.w)) when u.w > 0, which keeps
// FP32 precision for nearly straight chains.
__device__ __forceinline__ float angle_f(float ux, float uy, float uz, float wx, float wy,
                                         float wz, float ka) {
  const float lu = sqrtf(fmaf(ux, ux, fmaf(uy, uy, uz * uz)));
  const float lw = sqrtf(fmaf(wx, wx, fmaf(wy, wy, wz * wz)));
  if (!(lu > 1e-6f && lw > 1e-6f)) return 0.f;
  const float dot = fmaf(ux, wx, fmaf(uy, wy, uz * wz));
  const float ll = lu * lw;
  float omc;
  if (dot > 0.f) {
    const float cx = uy * wz - uz * wy, cy = uz * wx - ux * wz, cz = ux * wy - uy * wx;
    const float cr = fmaf(cx, cx, fmaf(cy, cy, cz * cz));
    omc = cr / (ll * (ll + dot));
  } else {
    omc = 1.f - dot / ll;
  }
  omc = fminf(fmaxf(omc, 0.f), 2.f);
  const float ang = 0.5f * omc;
  return ang * ang * ang * ka;
}

// The move loops' form of the smooth-level ΔE.  Per evaluation only the two
// bonds at the moved bead change length; the bonds (p-2,p-1) and (p+1,p+2)
// are shared by both evaluations, so each length is computed once per move
// (6 square roots instead of 16) and every bond length serves its bond term
// and both adjacent angles.  Square roots and quotients use the hardware
// v_sqrt_f32 / v_rcp_f32 (1 ulp each; the correctly rounded forms expand to
// ~10-15 VALU instructions apiece and were half of this loop's VALU work),
// and the angle's two quotient forms share one reciprocal, without a branch.
// Local window of bead p, unweighted (ParallelMonteCarloSmooth.cu:86-154):
// bonds (p-1,p), (p,p+1), then the angles at p-1, p, p+1.
__device__ __forceinline__ float len3(float x, float y, float z) {
  return __builtin_amdgcn_sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
}

// bond term of length d with rest length d0 (rd0 = 1/d0)
__device__ __forceinline__ float bond_r(float d, float d0, float rd0, float ks, float kq) {
  const float diff = (d - d0) * rd0;
  return diff * diff * (diff >= 0.f ? ks : kq);
}

// angle term of angle_f from bond vectors u, w with lengths lu, lw
__device__ __forceinline__ float angle_l(float ux, float uy, float uz, float lu, float wx, float wy,
                                         float wz, float lw, float ka) {
  const float dot = fmaf(ux, wx, fmaf(uy, wy, uz * wz));
  const float ll = lu * lw;
  const float cx = uy * wz - uz * wy, cy = uz * wx - ux * wz, cz = ux * wy - uy * wx;
  const float cr = fmaf(cx, cx, fmaf(cy, cy, cz * cz));
  const bool acute = dot > 0.f;
  // 1 - cos = |u x w|^2 / (ll (ll + u.w)) for u.w > 0, else (ll - u.w) / ll
  const float num = acute ? cr : ll - dot;
  const float den = acute ? ll * (ll + dot) : ll;
  const float omc = fminf(fmaxf(num * __builtin_amdgcn_rcpf(den), 0.f), 2.f);
  const float ang = 0.5f * omc;
  // short bonds: no angle term (a zero length leaves a NaN or inf above,
  // which the select drops)
  return (lu > 1e-6f && lw > 1e-6f) ? ang * ang * ang * ka : 0.f;
}

// ΔE = local(new) - local(old) of bead p; q* hold beads p-2..p+2 (index 2
// ignored; zero outside the chain), (px,py,pz) the old and (nx,ny,nz) the
// new centre.
__device__ __forceinline__ float smooth_delta(const DevProblem &P, int p, const float *qx,
                                              const float *qy, const float *qz, float px, float py,
                                              float pz, float nx, float ny, float nz, float d0m,
                                              float d0p) {
  const int n = P.n;
  const bool bm = p > 0, bp = p < n - 1;
  const bool am = p >= 2, ac = p >= 1 && p + 1 <= n - 1, ap = p + 2 <= n - 1;
  const float ux = qx[0] - qx[1], uy = qy[0] - qy[1], uz = qz[0] - qz[1];
  const float wx = qx[3] - qx[4], wy = qy[3] - qy[4], wz = qz[3] - qz[4];
  const float l0 = am ? len3(ux, uy, uz) : 0.f;
  const float l3 = ap ? len3(wx, wy, wz) : 0.f;
  const float rm = __builtin_amdgcn_rcpf(d0m), rp = __builtin_amdgcn_rcpf(d0p);
  float e[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const float cx = t ? nx : px, cy = t ? ny : py, cz = t ? nz : pz;
    const float v1x = qx[1] - cx, v1y = qy[1] - cy, v1z = qz[1] - cz;  // bead p-1 -> p
    const float v2x = cx - qx[3], v2y = cy - qy[3], v2z = cz - qz[3];  // bead p -> p+1
    const float l1 = len3(v1x, v1y, v1z), l2 = len3(v2x, v2y, v2z);
    float s = 0.f;
    if (bm) s += bond_r(l1, d0m, rm, P.ks, P.kq);
    if (bp) s += bond_r(l2, d0p, rp, P.ks, P.kq);
    if (am) s += angle_l(ux, uy, uz, l0, v1x, v1y, v1z, l1, P.ka);
    if (ac) s += angle_l(v1x, v1y, v1z, l1, v2x, v2y, v2z, l2, P.ka);
    if (ap) s += angle_l(v2x, v2y, v2z, l2, wx, wy, wz, l3, P.ka);
    e[t] = s;
  }
  return e[1] - e[0];
}

// Everything one smooth-level move reads before its decision, fetched ahead
// of time: the bead, the positions of beads p-2..p+2 (index 2 = p; zero
// outside the chain) and the two rest lengths of its bonds.
struct SmoothFetch {
  int p;
  float qx[5], qy[5], qz[5];
  float d0m, d0p;
};

// the 5-bead window's positions (outside the chain: 0)
template <class View>
__device__ __forceinline__ void smooth_fetch_pos(const DevProblem &P, const View &X, int p,
                                                 SmoothFetch &f) {
  const int n = P.n;
  f.p = p;
#pragma unroll
  for (int o = 0; o < 5; ++o) {
    const int idx = p - 2 + o;
    const bool in = idx >= 0 && idx < n;
    f.qx[o] = in ? X.get(idx, 0) : 0.f;
    f.qy[o] = in ? X.get(idx, 1) : 0.f;
    f.qz[o] = in ? X.get(idx, 2) : 0.f;
  }
}

// The bead picks of 64 consecutive smooth moves mw .. mw + 63 and their rest
// lengths, gathered one move per lane as ArcsWindow does for arc moves (the
// free-list entry and the two rest lengths were dependent scalar loads in
// every move's fetch).  dist_to_next holds n entries, so both clamped
// indices stay in it.
struct SmoothWindow {
  int p;
  float d0m, d0p;
};

__device__ __forceinline__ void smooth_window(const DevProblem &P, uint64_t seed, uint32_t group,
                                              uint32_t segment, uint32_t mw, SmoothWindow &w) {
  const int n = P.n;
  const uint32_t m = mw + static_cast<uint32_t>(threadIdx.x & (kLanes - 1));
  const U4 rb = draw(seed, m >> 2, group, segment, PMMC_LEVEL_SMOOTH, kSubBead);
  const int p = P.free_list[pick(pick_word(rb, m), static_cast<uint32_t>(P.nfree))];
  w.p = p;
  w.d0m = p > 0 ? fmaxf(P.dtn[max(p - 1, 0)], 1e-6f) : 1.f;
  w.d0p = p < n - 1 ? fmaxf(P.dtn[p], 1e-6f) : 1.f;
}

// smooth_fetch of move mw + k from its window
template <class View>
__device__ __forceinline__ void smooth_fetch_win(const DevProblem &P, const View &X,
                                                 const SmoothWindow &w, int k, SmoothFetch &f) {
  smooth_fetch_pos(P, X, __builtin_amdgcn_readlane(w.p, k), f);
  f.d0m = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(w.d0m), k));
  f.d0p = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(w.d0p), k));
}

// The fetched window holds positions from before the current move; when the
// current move changed bead q (per lane: ok), its copy takes the new position.
__device__ __forceinline__ void smooth_patch(SmoothFetch &f, int q, bool ok, float x, float y,
                                             float z) {
  const int o = q - (f.p - 2);
  if (o >= 0 && o < 5) {
#pragma unroll
    for (int u = 0; u < 5; ++u) {
      if (u == o) {
        f.qx[u] = ok ? x : f.qx[u];
        f.qy[u] = ok ? y : f.qy[u];
        f.qz[u] = ok ? z : f.qz[u];
      }
    }
  }
}

// ------------------------------------------------------------------ //
// milestone re-summation (FP32 terms, FP64 sum), per lane
// ------------------------------------------------------------------ //
template <class View>
__device__ double resum_heatmap(const DevProblem &P, const View &X, float r) {
  const int n = P.n;
  double tot = 0.0;
  const int lane = threadIdx.x & (kLanes - 1);
  for (int m = 0; m < n; ++m) {
    const float xm = X.get(m, 0), ym = X.get(m, 1), zm = X.get(m, 2);
    const float PMMC_AS_GLOBAL *wr = (const float PMMC_AS_GLOBAL *)P.W + static_cast<size_t>(m) * P.ldw;
    float part = 0.f;
    const int j0 = P.symmetric ? m + 1 : 0;
    for (int c = j0; c < n; c += kLanes) {
      const int cnt = min(kLanes, n - c);
      const float wv = lane < cnt ? wr[c + lane] : 0.f;
      for (int jj = 0; jj < cnt; ++jj) {
        const float w = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wv), jj));
        if (w == 0.f) continue;
        const int j = c + jj;
        const float t = fmaf(dist3(xm, ym, zm, X.get(j, 0), X.get(j, 1), X.get(j, 2)), w, -1.f);
        part = fmaf(t, t, part);
      }
    }
    tot += static_cast<double>(part);
  }
  if (P.symmetric) tot *= 2.0;
  if (r > 0.f)
    for (int m = 0; m < n; ++m) tot += sphere_penalty(X.get(m, 0), X.get(m, 1), X.get(m, 2), r);
  return tot;
}

// Arc + loop terms over the pair list (each pair once), eight pairs per
// round trip: the (wave-uniform) pair entries and then the 48 per-lane
// position loads are all issued before any is consumed.  FP32 partial sums of
// at most 64 pairs are added into the FP64 total.
template <class View>
__device__ double resum_arcs(const DevProblem &P, const View &X, float r, double offset) {
  const int n = P.n;
  const int np = P.npairs;
  double tot = offset;
  float part = 0.f;
  for (int q = 0; q < np; q += 8) {
    PairEntry pe[8];
    float xi[8], yi[8], zi[8], xj[8], yj[8], zj[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) pe[u] = load_pair(P.pairs, min(q + u, np - 1));
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = pe[u].i, j = pe[u].e.j & 0x7fffffff;
      xi[u] = X.get(i, 0);
      yi[u] = X.get(i, 1);
      zi[u] = X.get(i, 2);
      xj[u] = X.get(j, 0);
      yj[u] = X.get(j, 1);
      zj[u] = X.get(j, 2);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (q + u < np) {
        const float t = dist3(xi[u], yi[u], zi[u], xj[u], yj[u], zj[u]) - pe[u].e.c;
        part = fmaf(t * t, t >= 0.f ? pe[u].e.ka : pe[u].e.kb, part);
      }
    }
    if ((q & 63) == 56) {
      tot += static_cast<double>(part);
      part = 0.f;
    }
  }
  tot += static_cast<double>(part);
  if (r > 0.f)
    for (int m = 0; m < n; ++m) tot += sphere_penalty(X.get(m, 0), X.get(m, 1), X.get(m, 2), r);
  return tot;
}

template <class View>
__device__ double resum_smooth(const DevProblem &P, const View &X, float r, double offset) {
  const int n = P.n;
  double tot = offset;
  float part = 0.f;
  int cnt = 0;
  // a = bead i-1, b = bead i, c = bead i+1
  float ax = 0.f, ay = 0.f, az = 0.f;
  float bx = X.get(0, 0), by = X.get(0, 1), bz = X.get(0, 2);
  for (int i = 0; i + 1 < n; ++i) {
    const float cx = X.get(i + 1, 0), cy = X.get(i + 1, 1), cz = X.get(i + 1, 2);
    part += bond_f(dist3(bx, by, bz, cx, cy, cz), fmaxf(cst(P.dtn)[i], 1e-6f), P.ks, P.kq);
    if (i >= 1) part += angle_f(ax - bx, ay - by, az - bz, bx - cx, by - cy, bz - cz, P.ka);
    ax = bx;
    ay = by;
    az = bz;
    bx = cx;
    by = cy;
    bz = cz;
    if (++cnt == 256) {  // flush FP32 partials into the FP64 total
      tot += static_cast<double>(part);
      part = 0.f;
      cnt = 0;
    }
  }
  tot += static_cast<double>(part);
  if (r > 0.f)
    for (int m = 0; m < n; ++m) tot += sphere_penalty(X.get(m, 0), X.get(m, 1), X.get(m, 2), r);
  return tot;
}

}  // namespace pmmc
