// Equidistant model: Catmull-Rom positions at regularly spaced genomic
// positions along one chromosome's bead chain
// (HierarchicalChromosome::gpuHelper, HierarchicalChromosome.cu:1252-1449,
// over the kernel parallelFind3DSmoothPosition, :420-464).
//
// Output k sits at genomic position g = pos_start + k*resolution.  Before the
// first bead's end it is the first bead, from the last bead's start on the
// last bead; otherwise the segment (i-1, i) is the first bead i >= 1 with
// g < genomic_pos[i] (the beads are sorted by genomic position) and the point
// is the Catmull-Rom blend of beads i-2..i+1 at t = (g - gp[i-1]) /
// (gp[i] - gp[i-1]), with mirrored end points 2*p1 - p2 / 2*p2 - p1.
//
// One thread per output (grid-stride, every output written; the reference
// kernel returns after its first output, so outputs beyond one grid stay
// unwritten there).  The segment is found by binary search instead of the
// reference's linear scan; the arithmetic is the reference CPU path's, each
// product and sum rounded separately (no fused multiply-add), so results
// match it bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "ctx.hpp"
#include "pmmc_hip.h"

namespace pmmc {

struct SplineArgs {
  const int32_t *gp;   // [n] genomic positions, non-decreasing
  const float *xyz;    // [n*3]
  int32_t n;
  int32_t first_end;   // end of the first bead
  int32_t last_start;  // start of the last bead
  int32_t resolution;
  int64_t pos_start;
  int64_t n_out;
  float *out;          // [n_out*3]
};

__device__ __forceinline__ void load3(const float *xyz, int i, float &x, float &y, float &z) {
  x = xyz[3 * i];
  y = xyz[3 * i + 1];
  z = xyz[3 * i + 2];
}

__global__ __launch_bounds__(256) void k_spline(SplineArgs a) {
#pragma clang fp contract(off)
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < a.n_out; k += stride) {
    const int64_t g = a.pos_start + static_cast<int64_t>(a.resolution) * k;
    float rx, ry, rz;
    if (g < a.first_end) {
      load3(a.xyz, 0, rx, ry, rz);
    } else if (g >= a.last_start) {
      load3(a.xyz, a.n - 1, rx, ry, rz);
    } else {
      // first i with g < gp[i]
      int lo = 0, hi = a.n;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (g < a.gp[mid])
          hi = mid;
        else
          lo = mid + 1;
      }
      // the reference's scan only interpolates from its second bead on, so a
      // position before every bead uses segment (0, 1) (t < 0); past every
      // bead it keeps the last bead
      const int i = lo == 0 ? 1 : lo;
      if (i >= a.n) {
        load3(a.xyz, a.n - 1, rx, ry, rz);
      } else {
        float x0, y0, z0, x1, y1, z1, x2, y2, z2, x3, y3, z3;
        load3(a.xyz, i - 1, x1, y1, z1);
        load3(a.xyz, i, x2, y2, z2);
        if (i > 1) {
          load3(a.xyz, i - 2, x0, y0, z0);
        } else {
          x0 = 2.0f * x1 - x2;
          y0 = 2.0f * y1 - y2;
          z0 = 2.0f * z1 - z2;
        }
        if (i + 1 < a.n) {
          load3(a.xyz, i + 1, x3, y3, z3);
        } else {
          x3 = 2.0f * x2 - x1;
          y3 = 2.0f * y2 - y1;
          z3 = 2.0f * z2 - z1;
        }
        const float t = static_cast<float>(g - a.gp[i - 1]) / static_cast<float>(a.gp[i] - a.gp[i - 1]);
        const float t2 = t * t, t3 = t2 * t;
        const float b1 = 0.5f * (-t3 + 2.0f * t2 - t);
        const float b2 = 0.5f * (3.0f * t3 - 5.0f * t2 + 2.0f);
        const float b3 = 0.5f * (-3.0f * t3 + 4.0f * t2 + t);
        const float b4 = 0.5f * (t3 - t2);
        rx = x0 * b1 + x1 * b2 + x2 * b3 + x3 * b4;
        ry = y0 * b1 + y1 * b2 + y2 * b3 + y3 * b4;
        rz = z0 * b1 + z1 * b2 + z2 * b3 + z3 * b4;
      }
    }
    a.out[3 * k] = rx;
    a.out[3 * k + 1] = ry;
    a.out[3 * k + 2] = rz;
  }
}

}  // namespace pmmc

using namespace pmmc;

extern "C" int64_t pmmc_spline_count(int64_t pos_start, int64_t pos_stop, int32_t resolution_bp) {
  if (resolution_bp <= 0 || pos_stop < pos_start) return 0;
  return (pos_stop - pos_start + 1) / resolution_bp + 1;
}

extern "C" pmmc_status pmmc_spline_positions(pmmc_ctx *ctx, int32_t n_beads,
                                             const int32_t *genomic_pos, const float *xyz,
                                             int32_t first_end, int32_t last_start,
                                             int64_t pos_start, int32_t resolution_bp,
                                             int64_t n_out, float *out_xyz) {
  if (!ctx) return PMMC_ERR_INVALID;
  if (n_beads < 1 || !genomic_pos || !xyz || resolution_bp <= 0 || n_out < 0 || (n_out > 0 && !out_xyz))
    return fail(ctx, PMMC_ERR_INVALID, "spline: bad sizes or null arrays");
  for (int32_t i = 1; i < n_beads; ++i)
    if (genomic_pos[i] < genomic_pos[i - 1])
      return fail(ctx, PMMC_ERR_INVALID, "spline: beads must be sorted by genomic position");
  if (n_out == 0) return PMMC_OK;
  Sizer z;
  z.add<int32_t>(n_beads);
  z.add<float>(3 * static_cast<size_t>(n_beads));
  z.add<float>(3 * static_cast<size_t>(n_out));
  pmmc_status ps = arena_reserve(ctx, z.top + 4096);
  if (ps != PMMC_OK) return ps;
  Carver cv{ctx->arena, 0, ctx->arena_cap};
  SplineArgs a;
  int32_t *d_gp = cv.take<int32_t>(n_beads);
  float *d_xyz = cv.take<float>(3 * static_cast<size_t>(n_beads));
  a.out = cv.take<float>(3 * static_cast<size_t>(n_out));
  a.gp = d_gp;
  a.xyz = d_xyz;
  a.n = n_beads;
  a.first_end = first_end;
  a.last_start = last_start;
  a.resolution = resolution_bp;
  a.pos_start = pos_start;
  a.n_out = n_out;
  hipStream_t s = ctx->stream;
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_gp, genomic_pos, sizeof(int32_t) * n_beads, hipMemcpyHostToDevice, s));
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(d_xyz, xyz, sizeof(float) * 3 * n_beads, hipMemcpyHostToDevice, s));
  const int64_t blocks = std::min<int64_t>((n_out + 255) / 256, 65536);
  hipLaunchKernelGGL(k_spline, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, a);
  PMMC_HIP_TRY(ctx, hipGetLastError());
  PMMC_HIP_TRY(ctx, hipMemcpyAsync(out_xyz, a.out, sizeof(float) * 3 * n_out, hipMemcpyDeviceToHost, s));
  PMMC_HIP_TRY(ctx, hipStreamSynchronize(s));
  return PMMC_OK;
}
