// device.hpp -- device-side primitives shared by the annealing kernels.
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), keyed by the 64-bit run seed, with the
// 128-bit counter {move, chain, segment, level<<28 | sub}.  The curand XORWOW
// generator of the reference (curand_init(seed, tid, 0), e.g.
// ParallelMonteCarloArcs.cu:385-388) is replaced by this stateless stream so
// that a chain's draws depend only on its logical coordinates, never on the
// thread or slot that runs it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmmc {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

constexpr uint32_t kSubChain = 0;  // {dx, dy, dz, U} of one move of one chain
constexpr uint32_t kSubBead = 1;   // bead picks of 4 moves of one pick group

struct U4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                      uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t a = static_cast<uint64_t>(kPhiloxM0) * c0;
    const uint64_t b = static_cast<uint64_t>(kPhiloxM1) * c2;
    const uint32_t n0 = static_cast<uint32_t>(b >> 32) ^ c1 ^ k0;
    const uint32_t n2 = static_cast<uint32_t>(a >> 32) ^ c3 ^ k1;
    c1 = static_cast<uint32_t>(b);
    c3 = static_cast<uint32_t>(a);
    c0 = n0;
    c2 = n2;
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  return U4{c0, c1, c2, c3};
}

__host__ __device__ __forceinline__ U4 draw(uint64_t seed, uint32_t ctr0, uint32_t ctr1,
                                             uint32_t segment, uint32_t level, uint32_t sub) {
  return philox4x32_10(ctr0, ctr1, segment, (level << 28) | (sub & 0x0FFFFFFFu),
                       static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
}

// [0, 1) with 24 random bits; exact in FP32.
__host__ __device__ __forceinline__ float u01(uint32_t x) {
  return static_cast<float>(x >> 8) * 0x1p-24f;
}

// multiply-shift index in [0, m)
__host__ __device__ __forceinline__ uint32_t pick(uint32_t x, uint32_t m) {
  return static_cast<uint32_t>((static_cast<uint64_t>(x) * m) >> 32);
}

// displacement component (2u - 1) * step; 2u - 1 is exact, one rounding.
// The product and the position update are kept as separate IEEE operations
// (no fused multiply-add), exactly as the reference's vector3 += does.
// (__fmul_rn/__fadd_rn alone do not stop the backend from fusing them.)
__device__ __forceinline__ float disp(uint32_t x, float step) {
#pragma clang fp contract(off)
  const float t = 2.0f * u01(x) - 1.0f;
  return t * step;
}

__device__ __forceinline__ float move_coord(float x, float d) {
#pragma clang fp contract(off)
  return x + d;
}

// Rounded as written (no contraction): the arc kernels inline it in
// different surroundings and must give the same bits.
__device__ __forceinline__ float sphere_penalty(float x, float y, float z, float r) {
#pragma clang fp contract(off)
  if (r <= 0.0f) return 0.0f;
  const float d = sqrtf(x * x + y * y + z * z);
  if (d <= r) return 0.0f;
  const float o = d - r;
  return o * o;
}

__device__ __forceinline__ double sphere_penalty_d(float x, float y, float z, float r) {
  if (r <= 0.0f) return 0.0;
  const double d = sqrt((double)x * x + (double)y * y + (double)z * z);
  if (d <= r) return 0.0;
  const double o = d - r;
  return o * o;
}

// FP32 distance with the hardware square root (v_sqrt_f32, 1 ulp).  The
// correctly rounded sqrtf expands to ~15 VALU instructions around it, which
// made the square roots the largest VALU cost of the move loops; the move
// deltas keep their 1e-5 relative bound against the FP64 energies.
__device__ __forceinline__ float dist3(float ax, float ay, float az, float bx, float by,
                                       float bz) {
  const float dx = ax - bx, dy = ay - by, dz = az - bz;
  return __builtin_amdgcn_sqrtf(fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
}

__device__ __forceinline__ double dist3d(float ax, float ay, float az, float bx, float by,
                                         float bz) {
  const double dx = (double)ax - bx, dy = (double)ay - by, dz = (double)az - bz;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// Wavefront sum of a double, fixed reduction order (deterministic).
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace pmmc
