// acosf_fdlibm.hpp -- single-precision arc cosine, the published fdlibm
// method (Sun Microsystems' e_acos.c, float form as glibc's flt-32/e_acosf.c
// computes it): |x| < 0.5 -> pi/2 - (x + x*R(x^2)); x < -0.5 -> pi - 2(s +
// s*R(z)) with z = (1 + x)/2, s = sqrt(z); x > 0.5 -> 2(df + (c + s*R(z)))
// with z = (1 - x)/2, df = s with its low 12 bits cleared and
// c = (z - df^2)/(s + df); R = P(z)/Q(z), a degree-(6, 4) rational fit.
//
// The reference's host code calls libm's acosf (the angular terms of
// LooperSolver::stressRefineAnchors, LooperSolver.hpp:5826-5858); this form
// returns the same bits as glibc 2.35's acosf for every float in [-1, 1]
// (checked exhaustively against libm by tests/test_stress.py), so the same
// source compiled for the GPU and for the host gives the reference's angles
// on both.  Every operation is a single IEEE-rounded float operation: the
// including file must not contract products and sums into fused
// multiply-adds, and sqrt and division must be correctly rounded (HIP's
// default for float).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define PMMC_HD __host__ __device__
#else
#define PMMC_HD
#endif
#if defined(__clang__)
#define PMMC_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define PMMC_NO_CONTRACT
#endif

namespace pmmc {

PMMC_HD inline float f32_of(uint32_t u) { return __builtin_bit_cast(float, u); }
PMMC_HD inline uint32_t bits_of(float f) { return __builtin_bit_cast(uint32_t, f); }

PMMC_HD inline float acosf_fdlibm(float x) {
  PMMC_NO_CONTRACT
  // coefficients as float bit patterns
  const float pi = f32_of(0x40490fdau), pio2_hi = f32_of(0x3fc90fdau), pio2_lo = f32_of(0x33a22168u);
  const float p0 = f32_of(0x3e2aaaabu), p1 = f32_of(0xbea6b090u), p2 = f32_of(0x3e4e0aa8u),
              p3 = f32_of(0xbd241146u), p4 = f32_of(0x3a4f7f04u), p5 = f32_of(0x3811ef08u);
  const float q1 = f32_of(0xc019d139u), q2 = f32_of(0x4001572du), q3 = f32_of(0xbf303361u),
              q4 = f32_of(0x3d9dc62eu);
  const uint32_t hx = bits_of(x), ix = hx & 0x7fffffffu;
  const bool neg = (hx >> 31) != 0;
  if (ix == 0x3f800000u) return neg ? pi + 2.0f * pio2_lo : 0.0f;  // |x| = 1
  if (ix > 0x3f800000u) return (x - x) / (x - x);                  // |x| > 1, NaN
  if (ix < 0x3f000000u) {                                          // |x| < 0.5
    if (ix <= 0x32800000u) return pio2_hi + pio2_lo;               // |x| <= 2^-26
    const float z = x * x;
    const float p = z * (p0 + z * (p1 + z * (p2 + z * (p3 + z * (p4 + z * p5)))));
    const float q = 1.0f + z * (q1 + z * (q2 + z * (q3 + z * q4)));
    const float r = p / q;
    return pio2_hi - (x - (pio2_lo - x * r));
  }
  if (neg) {  // x <= -0.5
    const float z = (1.0f + x) * 0.5f;
    const float p = z * (p0 + z * (p1 + z * (p2 + z * (p3 + z * (p4 + z * p5)))));
    const float q = 1.0f + z * (q1 + z * (q2 + z * (q3 + z * q4)));
    const float s = std::sqrt(z);
    const float r = p / q;
    const float w = r * s - pio2_lo;
    return pi - 2.0f * (s + w);
  }
  // x >= 0.5
  const float z = (1.0f - x) * 0.5f;
  const float s = std::sqrt(z);
  const float df = f32_of(bits_of(s) & 0xfffff000u);
  const float c = (z - df * df) / (s + df);
  const float p = z * (p0 + z * (p1 + z * (p2 + z * (p3 + z * (p4 + z * p5)))));
  const float q = 1.0f + z * (q1 + z * (q2 + z * (q3 + z * q4)));
  const float r = p / q;
  const float w = r * s + c;
  return 2.0f * (df + w);
}

}  // namespace pmmc
