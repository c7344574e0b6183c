// probe.hip -- bandwidth probes for the annealers' access patterns.
//
// The roofline of the move kernels is HBM, but not a streaming HBM: a move of
// an ensemble reads a few random 768-byte chunks (one bead's three 256-byte
// SoA rows) and writes one of them back.  These kernels measure what the GPU
// sustains for exactly that pattern with no arithmetic, so a kernel's
// throughput can be set against the ceiling of its own pattern as well as
// against the nominal peak:
//   k_probe_moves<CB, MODE>: per move, `reads` random windows of CB consecutive
//     beads (CB x 768 B: SoA rows for the arc move, CB 1; each lane's
//     coordinates together for the smooth move, CB 5, as the two levels'
//     kernels lay out positions), then the first window's first bead written back by
//     no lane (MODE 0), about half of the lanes (1), every lane (2) or one
//     lane in about every other move (3: a low acceptance rate, as at the
//     smooth level); two moves in flight per wavefront, as the pipelined move
//     loops keep.
//   k_probe_stream: a grid-stride float4 read of the whole buffer.
//   k_probe_lds: conflict-free ds_read_b128 loops over 64 KB of LDS per
//     workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ctx.hpp"

namespace pmmc {

namespace {

constexpr int kProbeLanes = 64;
// reads per move: up to 8 single beads (the arc move: bead + partners), or
// one 5-bead window (the smooth move)
constexpr int max_reads(int cb) { return cb == 1 ? 8 : 1; }

__device__ __forceinline__ uint32_t probe_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// buf holds n_beads beads of 3 x 64 floats; windows start at bead
// hash % (n_beads - CB + 1), so every access stays inside the buffer.
template <int CB, int MODE>
__global__ __launch_bounds__(64) void k_probe_moves(float *__restrict__ buf, uint32_t n_starts,
                                                   int reads, int moves, float *__restrict__ sink) {
  constexpr int MR = max_reads(CB);
  const int lane = threadIdx.x;
  const uint32_t seed = blockIdx.x * 7919u + 17u;
  float acc = 0.f;
  for (int it = 0; it < moves; it += 2) {
    float v[2][MR][CB * 3];
    uint32_t first[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
#pragma unroll
      for (int r = 0; r < MR; ++r) {
        if (r < reads) {
          const uint32_t b = probe_hash(seed * 13u + static_cast<uint32_t>(it + u) * 8u + r) % n_starts;
          if (r == 0) first[u] = b;
          const float *w = buf + static_cast<size_t>(b) * 3 * kProbeLanes;
          if constexpr (CB == 1) {
#pragma unroll
            for (int k = 0; k < 3; ++k) v[u][r][k] = w[k * kProbeLanes + lane];
          } else {
#pragma unroll
            for (int k = 0; k < CB * 3; ++k) v[u][r][k] = w[(k / 3) * 3 * kProbeLanes + 3 * lane + k % 3];
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      float a[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < MR; ++r)
        if (r < reads)
#pragma unroll
          for (int k = 0; k < CB * 3; ++k) a[k % 3] += v[u][r][k];
      float *w = buf + static_cast<size_t>(first[u]) * 3 * kProbeLanes;
      // every move's loads feed acc, so the compiler cannot sink them into
      // the write-back branch (where they would be issued only when a lane
      // writes, and waited for at once)
      acc += a[0] + a[1] + a[2];
      if constexpr (MODE != 0) {
        const bool take = MODE == 2 ||
                          (MODE == 1 && (probe_hash(seed * 31u + it + u + lane * 977u) & 1u)) ||
                          (MODE == 3 && lane == (probe_hash(seed + it + u) & 127u));
        // the written value equals what was read (the sums are scaled to 0):
        // the buffer keeps its contents
        // as the kernels write back: non-temporal stores over the arc
        // level's SoA rows, plain stores of the smooth level's lane triples
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if (take) {
            if constexpr (CB == 1)
              __builtin_nontemporal_store(fmaf(a[k], 0.f, v[u][0][k]), &w[k * kProbeLanes + lane]);
            else
              w[3 * lane + k] = fmaf(a[k], 0.f, v[u][0][k]);
          }
      }
    }
  }
  if (acc == 12345.f) sink[blockIdx.x & 65535] = acc;  // never taken: the buffer holds zeros
}

__global__ void k_probe_stream(const float4 *__restrict__ buf, size_t n4, float *__restrict__ sink) {
  float acc = 0.f;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n4;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const float4 v = buf[i];
    acc += v.x + v.y + v.z + v.w;
  }
  if (acc == 12345.f) sink[blockIdx.x] = acc;
}

// 256 threads read 64 KB of LDS, 16 B per lane per access at consecutive
// addresses (conflict-free), `passes` times.
__global__ __launch_bounds__(256) void k_probe_lds(int passes, float *__restrict__ sink) {
  __shared__ float4 s[4096];
  for (int i = threadIdx.x; i < 4096; i += 256) s[i] = make_float4(i, 0.f, 0.f, 0.f);
  __syncthreads();
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int p = 0; p < passes; ++p) {
#pragma unroll 4
    for (int i = threadIdx.x; i < 4096; i += 256) {
      const float4 v = s[(i + p) & 4095];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
  }
  if (acc.x + acc.y + acc.z + acc.w == 12345.f) sink[blockIdx.x] = acc.x;
}

template <int CB>
hipError_t launch_moves(int mode, dim3 g, hipStream_t st, float *buf, uint32_t n_starts, int reads,
                        int moves, float *sink) {
  if (mode == 0)
    hipLaunchKernelGGL((k_probe_moves<CB, 0>), g, dim3(kProbeLanes), 0, st, buf, n_starts, reads, moves, sink);
  else if (mode == 1)
    hipLaunchKernelGGL((k_probe_moves<CB, 1>), g, dim3(kProbeLanes), 0, st, buf, n_starts, reads, moves, sink);
  else if (mode == 2)
    hipLaunchKernelGGL((k_probe_moves<CB, 2>), g, dim3(kProbeLanes), 0, st, buf, n_starts, reads, moves, sink);
  else
    hipLaunchKernelGGL((k_probe_moves<CB, 3>), g, dim3(kProbeLanes), 0, st, buf, n_starts, reads, moves, sink);
  return hipGetLastError();
}

}  // namespace

}  // namespace pmmc

using namespace pmmc;

extern "C" pmmc_status pmmc_probe_move_pattern(pmmc_ctx *ctx, size_t bytes, int32_t waves,
                                               int32_t moves_per_wave, int32_t reads,
                                               int32_t chunk_beads, int32_t write_mode,
                                               pmmc_probe_result *out) {
  if (!ctx || !out) return fail(ctx, PMMC_ERR_INVALID, "probe: null argument");
  if (waves < 1 || waves > (1 << 20) || moves_per_wave < 2 || moves_per_wave % 2 != 0 ||
      (chunk_beads != 1 && chunk_beads != 5) || reads < 1 || reads > max_reads(chunk_beads) ||
      write_mode < 0 || write_mode > 3)
    return fail(ctx, PMMC_ERR_INVALID,
                "probe: waves in [1, 2^20], even moves_per_wave >= 2, chunk_beads 1 (reads 1..8) "
                "or 5 (reads 1), write_mode 0..3");
  const size_t bead_bytes = 3 * kProbeLanes * sizeof(float);
  const size_t n_beads = bytes / bead_bytes;
  if (n_beads < 16 || n_beads > 0xFFFFFFFFull)
    return fail(ctx, PMMC_ERR_INVALID, "probe: bytes must hold 16 to 2^32 beads of 768 B");
  const size_t sink_bytes = sizeof(float) * 65536;
  pmmc_status ps = arena_reserve(ctx, n_beads * bead_bytes + sink_bytes + 256);
  if (ps != PMMC_OK) return ps;
  float *buf = reinterpret_cast<float *>(ctx->arena);
  float *sink = reinterpret_cast<float *>(ctx->arena + n_beads * bead_bytes);
  hipStream_t st = ctx->stream;
  PMMC_HIP_TRY(ctx, hipMemsetAsync(buf, 0, n_beads * bead_bytes, st));
  const uint32_t n_starts = static_cast<uint32_t>(n_beads - chunk_beads + 1);
  auto moves = [&]() -> hipError_t {
    return chunk_beads == 1
               ? launch_moves<1>(write_mode, dim3(waves), st, buf, n_starts, reads, moves_per_wave, sink)
               : launch_moves<5>(write_mode, dim3(waves), st, buf, n_starts, reads, moves_per_wave, sink);
  };
  float ms = 0.f;
  // warm-up, then the timed launch of the move pattern
  PMMC_HIP_TRY(ctx, moves());
  PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
  PMMC_HIP_TRY(ctx, moves());
  PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
  PMMC_HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
  PMMC_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  const double n_moves = static_cast<double>(waves) * moves_per_wave;
  const double rd = n_moves * reads * chunk_beads * bead_bytes;
  const double wr_frac[4] = {0.0, 0.5, 1.0, 0.5 / kProbeLanes};
  const double wr = n_moves * bead_bytes * wr_frac[write_mode];
  out->ms = ms;
  out->moves_per_s = n_moves * kProbeLanes / (ms * 1e-3);
  out->read_gbps = rd / (ms * 1e-3) / 1e9;
  out->write_gbps = wr / (ms * 1e-3) / 1e9;
  // streaming read of the same buffer
  const size_t n4 = n_beads * bead_bytes / 16;
  hipLaunchKernelGGL(k_probe_stream, dim3(4096), dim3(256), 0, st, reinterpret_cast<const float4 *>(buf), n4, sink);
  PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
  hipLaunchKernelGGL(k_probe_stream, dim3(4096), dim3(256), 0, st, reinterpret_cast<const float4 *>(buf), n4, sink);
  PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
  PMMC_HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
  PMMC_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  out->stream_read_gbps = static_cast<double>(n4) * 16 / (ms * 1e-3) / 1e9;
  // LDS: 8 workgroups of 256 threads per CU-sized slice of the grid
  const int lds_groups = 256 * 8, passes = 256;
  hipLaunchKernelGGL(k_probe_lds, dim3(lds_groups), dim3(256), 0, st, passes, sink);
  PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
  hipLaunchKernelGGL(k_probe_lds, dim3(lds_groups), dim3(256), 0, st, passes, sink);
  PMMC_HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
  PMMC_HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
  PMMC_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  out->lds_gbps = static_cast<double>(lds_groups) * passes * 4096 * 16 / (ms * 1e-3) / 1e9;
  return PMMC_OK;
}
