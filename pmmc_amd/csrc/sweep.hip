// sweep.hip -- the reference GPU heatmap annealer's algorithm
// (MonteCarloHeatmapKernel, ParallelMonteCarloHeatmap.cu:217-363) as a
// race-free synchronous sweep.
//
// The reference gives every bead a warp of 32 proposers that each run 512
// Metropolis proposals on that bead against the (concurrently changing)
// global positions, then lets the lowest-scoring lanes race to write back.
// Here every iteration works on a snapshot: workgroup (bead b, replica r)
// holds the snapshot and the heatmap column of b in LDS, its 64 lanes are
// 64 proposers, the wave's lowest final score (ties: lowest lane) becomes the
// replica's candidate, a second kernel picks the best replica per bead
// (ties: lowest replica) and accepts it when it lowers b's local score, and
// a third re-sums the full score in FP64 for the stop test.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "device.hpp"
#include "engine_host.hpp"

namespace pmmc {

struct SweepArgs {
  int n, ldw, replicas, inner, proposers;
  uint64_t seed;
  uint32_t segment, iter;
  double T;
  float step, tjs, tjc, sphere_radius;
  int use2d;
  const float *W;        // [n*n] W[b*n+j] = valid(j,b) ? 1/h[j][b] : 0
  const float *Hc;       // [n*n] h or 0
  const uint8_t *fixed;  // [n]
  const float *snap;     // [n*3]
  float *pos;            // [n*3]
  float *cand;           // [n*replicas*4]
  float *sprev;          // [n]
  double *energy;        // [1]
};

__global__ __launch_bounds__(64) void k_sweep_propose(SweepArgs a) {
  extern __shared__ float sm[];
  float *sx = sm, *sy = sm + a.n, *sz = sm + 2 * a.n, *w = sm + 3 * a.n;
  const int b = blockIdx.x / a.replicas, r = blockIdx.x % a.replicas;
  const int lane = threadIdx.x;
  if (a.fixed[b]) return;
  for (int j = lane; j < a.n; j += 64) {
    sx[j] = a.snap[3 * j];
    sy[j] = a.snap[3 * j + 1];
    sz[j] = a.snap[3 * j + 2];
    w[j] = a.W[static_cast<size_t>(b) * a.ldw + j];
  }
  __syncthreads();
  float x = sx[b], y = sy[b], z = sz[b];
  float sp = a.sprev[b];
  const float Tf = static_cast<float>(a.T);
  const uint32_t chain = static_cast<uint32_t>(b) * a.proposers + r * 64 + lane;
  for (int k = 0; k < a.inner; ++k) {
    const U4 rc = draw(a.seed, a.iter * static_cast<uint32_t>(a.inner) + k, chain, a.segment,
                       kLevelSweep, kSubChain);
    const float nx = move_coord(x, disp(rc.x, a.step)), ny = move_coord(y, disp(rc.y, a.step));
    const float nz = a.use2d ? z : move_coord(z, disp(rc.z, a.step));
    float s0 = 0.f, s1 = 0.f;
    int j = 0;
    for (; j + 1 < a.n; j += 2) {
      const float w0 = w[j], w1 = w[j + 1];
      const float t0 = fmaf(dist3(nx, ny, nz, sx[j], sy[j], sz[j]), w0, -1.f);
      const float t1 = fmaf(dist3(nx, ny, nz, sx[j + 1], sy[j + 1], sz[j + 1]), w1, -1.f);
      s0 = w0 != 0.f ? fmaf(t0, t0, s0) : s0;
      s1 = w1 != 0.f ? fmaf(t1, t1, s1) : s1;
    }
    if (j < a.n) {
      const float t0 = fmaf(dist3(nx, ny, nz, sx[j], sy[j], sz[j]), w[j], -1.f);
      s0 = w[j] != 0.f ? fmaf(t0, t0, s0) : s0;
    }
    const float sn = s0 + s1 + sphere_penalty(nx, ny, nz, a.sphere_radius);
    bool ok = sn <= sp;
    if (!ok && a.T > 0.0) ok = u01(rc.w) < a.tjs * __expf(-a.tjc * (sn / sp) / Tf);
    if (ok) {
      x = nx;
      y = ny;
      z = nz;
      sp = sn;
    }
  }
  // wave argmin of (sp, lane)
  float bs = sp;
  int bl = lane;
  for (int off = 32; off > 0; off >>= 1) {
    const float os = __shfl_xor(bs, off, 64);
    const int ol = __shfl_xor(bl, off, 64);
    if (os < bs || (os == bs && ol < bl)) {
      bs = os;
      bl = ol;
    }
  }
  if (lane == bl) {
    float *c = a.cand + (static_cast<size_t>(b) * a.replicas + r) * 4;
    c[0] = sp;
    c[1] = x;
    c[2] = y;
    c[3] = z;
  }
}

__device__ double sweep_local64(const SweepArgs &a, int b, float x, float y, float z) {
  double e = 0.0;
  for (int j = 0; j < a.n; ++j) {
    const float h = a.Hc[static_cast<size_t>(b) * a.ldw + j];
    if (h == 0.f) continue;
    const double c = dist3d(x, y, z, a.snap[3 * j], a.snap[3 * j + 1], a.snap[3 * j + 2]) / h - 1.0;
    e += c * c;
  }
  return e;
}

__global__ __launch_bounds__(64) void k_sweep_pick(SweepArgs a) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= a.n) return;
  const float sx = a.snap[3 * b], sy = a.snap[3 * b + 1], sz = a.snap[3 * b + 2];
  a.pos[3 * b] = sx;
  a.pos[3 * b + 1] = sy;
  a.pos[3 * b + 2] = sz;
  if (a.fixed[b]) return;
  const float *c = a.cand + static_cast<size_t>(b) * a.replicas * 4;
  int br = 0;
  for (int r = 1; r < a.replicas; ++r)
    if (c[4 * r] < c[4 * br]) br = r;
  const float *bc = c + 4 * br;
  const double lc = sweep_local64(a, b, bc[1], bc[2], bc[3]);
  const double lg = sweep_local64(a, b, sx, sy, sz);
  if (lc < lg) {
    a.pos[3 * b] = bc[1];
    a.pos[3 * b + 1] = bc[2];
    a.pos[3 * b + 2] = bc[3];
  }
  a.sprev[b] = bc[0];
}

// full score (FP64) of a.pos, one workgroup of 256
__global__ __launch_bounds__(256) void k_sweep_energy(SweepArgs a) {
  __shared__ double red[256];
  double s = 0.0;
  for (int m = threadIdx.x; m < a.n; m += 256)
    for (int j = 0; j < a.n; ++j) {
      const float h = a.Hc[static_cast<size_t>(m) * a.ldw + j];
      if (h == 0.f) continue;
      const double c =
          dist3d(a.pos[3 * m], a.pos[3 * m + 1], a.pos[3 * m + 2], a.pos[3 * j], a.pos[3 * j + 1],
                 a.pos[3 * j + 2]) / h - 1.0;
      s += c * c;
    }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *a.energy = red[0];
}

__global__ __launch_bounds__(64) void k_fill_f(float *p, int n, float v) {
  for (int i = blockIdx.x * 64 + threadIdx.x; i < n; i += gridDim.x * 64) p[i] = v;
}

}  // namespace pmmc

using namespace pmmc;

extern "C" pmmc_status pmmc_heatmap_sweep(pmmc_ctx *ctx, const pmmc_heatmap_problem *p,
                                          const pmmc_sweep_schedule *s, uint64_t seed,
                                          uint32_t segment, float *pos_inout,
                                          pmmc_anneal_stats *stats) {
  if (!ctx || !p || !s || !pos_inout) return fail(ctx, PMMC_ERR_INVALID, "sweep: null argument");
  if (s->inner <= 0 || s->proposers <= 0 || s->proposers % 64 != 0 || s->max_iters < 0 ||
      !(s->dt > 0.0))
    return fail(ctx, PMMC_ERR_INVALID, "sweep: inner > 0, proposers a multiple of 64, dt > 0");
  pmmc_schedule dummy;
  std::memset(&dummy, 0, sizeof(dummy));
  dummy.T0 = 1.0;
  dummy.dt = 1.0;
  dummy.block = 1;
  dummy.milestone_period = 1;
  pmmc_run run{seed ? seed : ctx->default_seed, segment, 0, 64};
  HostProblem h;
  pmmc_status ps = prepare_problem(ctx, PMMC_LEVEL_HEATMAP, p, dummy, run, h);
  if (ps != PMMC_OK) return ps;
  const int n = h.n;
  const size_t lds = static_cast<size_t>(n) * 4 * sizeof(float);
  if (lds > 64 * 1024) return fail(ctx, PMMC_ERR_INVALID, "sweep: n too large for LDS staging (max 4096)");
  const int R = s->proposers / 64;
  Sizer z;
  z.add<float>(h.W.size());
  z.add<float>(h.Hc.size());
  z.add<uint8_t>(n);
  z.add<float>(3 * n);
  z.add<float>(3 * n);
  z.add<float>(static_cast<size_t>(n) * R * 4);
  z.add<float>(n);
  z.add<double>(1);
  ps = arena_reserve(ctx, z.top + 4096);
  if (ps != PMMC_OK) return ps;
  Carver cv{ctx->arena, 0, ctx->arena_cap};
  hipStream_t st = ctx->stream;
  SweepArgs a;
  a.n = n;
  a.ldw = h.ldw;
  a.replicas = R;
  a.inner = s->inner;
  a.proposers = s->proposers;
  a.seed = h.seed;
  a.segment = segment;
  a.tjs = s->tjs;
  a.tjc = s->tjc;
  a.sphere_radius = s->sphere_radius;
  a.use2d = s->use2d;
  float *W = cv.take<float>(h.W.size());
  float *Hc = cv.take<float>(h.Hc.size());
  uint8_t *fx = cv.take<uint8_t>(n);
  float *snap = cv.take<float>(3 * n);
  float *pos = cv.take<float>(3 * n);
  float *cand = cv.take<float>(static_cast<size_t>(n) * R * 4);
  float *sprev = cv.take<float>(n);
  double *energy = cv.take<double>(1);
  a.W = W;
  a.Hc = Hc;
  a.fixed = fx;
  a.snap = snap;
  a.pos = pos;
  a.cand = cand;
  a.sprev = sprev;
  a.energy = energy;
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(W, h.W.data(), sizeof(float) * h.W.size(), hipMemcpyHostToDevice, st));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(Hc, h.Hc.data(), sizeof(float) * h.Hc.size(), hipMemcpyHostToDevice, st));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(fx, h.fixed.data(), n, hipMemcpyHostToDevice, st));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(pos, pos_inout, sizeof(float) * 3 * n, hipMemcpyHostToDevice, st));
  auto full_energy = [&](double &E) -> pmmc_status {
    hipLaunchKernelGGL(k_sweep_energy, dim3(1), dim3(256), 0, st, a);
    PMMC_HIP_TRY(ctx, hipGetLastError());
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(&E, energy, sizeof(double), hipMemcpyDeviceToHost, st));
    PMMC_HIP_TRY(ctx, hipStreamSynchronize(st));
    return PMMC_OK;
  };
  double E = 0.0;
  ps = full_energy(E);
  if (ps != PMMC_OK) return ps;
  const float E0f = static_cast<float>(E);
  hipLaunchKernelGGL(k_fill_f, dim3(std::max(1, std::min(64, (n + 63) / 64))), dim3(64), 0, st, sprev, n, E0f);
  double T = s->T0, ms = E;
  float step = s->step;
  int misses = 0, reason = 0, it = 0;
  int n_free = 0;
  for (int b = 0; b < n; ++b) n_free += h.fixed[b] ? 0 : 1;
  PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
  if (!(E == E) || E < 1e-6) reason = PMMC_STOP_GUARD;
  while (reason == 0) {
    if (it >= s->max_iters) {
      reason = PMMC_STOP_CAP;
      break;
    }
    a.iter = static_cast<uint32_t>(it);
    a.T = T;
    a.step = step;
    PMMC_HIP_TRY(ctx, hipMemcpyAsync(snap, pos, sizeof(float) * 3 * n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_sweep_propose, dim3(n * R), dim3(64), lds, st, a);
    hipLaunchKernelGGL(k_sweep_pick, dim3((n + 63) / 64), dim3(64), 0, st, a);
    PMMC_HIP_TRY(ctx, hipGetLastError());
    T *= s->dt;
    step *= s->step_decay;
    ++it;
    ps = full_energy(E);
    if (ps != PMMC_OK) return ps;
    if (!std::isfinite(E)) {
      reason = PMMC_STOP_NONFINITE;
      break;
    }
    if (E > s->improvement * ms) ++misses;
    if (E < s->small_energy) {
      reason = PMMC_STOP_SMALL;
      break;
    }
    if (misses >= s->fails_threshold) {
      reason = PMMC_STOP_RULE;
      break;
    }
    ms = E;
  }
  PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(pos_inout, pos, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, st));
  PMMC_HIP_TRY(ctx, hipStreamSynchronize(st));
  if (stats) {
    float ms_k = 0.f;
    (void)hipEventElapsedTime(&ms_k, ctx->ev0, ctx->ev1);
    std::memset(stats, 0, sizeof(*stats));
    stats->energy = E;
    stats->energy_resummed = E;
    stats->best_energy = E;
    stats->moves = static_cast<int64_t>(it) * n_free * s->proposers * s->inner;
    stats->best_chain = -1;
    stats->n_chains = n_free * s->proposers;
    stats->milestones = it;
    stats->stop_reason = reason;
    stats->kernel_ms = ms_k;
  }
  return PMMC_OK;
}
