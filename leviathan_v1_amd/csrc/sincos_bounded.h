// sincos_bounded.h -- sine and cosine of a bounded argument, for the edge
// pass of the step.
//
// The edge pass takes sin and cos of the wrapped phase difference
// wrap(theta_j(t - d) - theta_i), which lies in [-pi, pi] (the reference uses
// the fast intrinsic __sincosf there, leviathan_kernel.cu:150-152).  A general
// sincosf carries a large-argument (Payne-Hanek) reduction that the compiler
// evaluates without a branch, ~150 vector instructions per call, which made
// the row pass compute-bound; this routine costs ~25.
//
// Method: Cody-Waite reduction x = k pi/2 + r with pi/2 split into three
// floats (pi/2 = P1 + P2 + P3 to 2^-76; every product fused, so r carries
// about one rounding error even next to the zeros of sin and cos, where
// |r| is ~1e-8), |r| <= pi/4; sin r and cos r from the polynomials
//   sin r = r + r^3 (S1 + S2 r^2 + S3 r^4)
//   cos r = 1 + r^2 (C1 + C2 r^2 + C3 r^4 + C4 r^6)
// whose coefficients were fitted for minimum maximum relative error on
// |r| <= pi/4 by Lawson's iteratively reweighted least squares in double
// precision (fit_sincos.py, next to this file) and rounded to float; then the
// quadrant k mod 4 selects and signs the results.  Error: within 2 ulp of the
// correctly rounded sin and cos for |x| <= kSinCosBoundedMax (checked by the
// CPU tests over a dense sweep); sin(-0) returns +0.  The caller handles |x|
// beyond that bound, and NaN, with the general routine.
//
// LEV_HD must be defined by the including file: the device build defines it
// as a host+device inline function, the CPU check as a plain inline one.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

namespace lev {

constexpr float kSinCosBoundedMax = 8.0f;

// x with its sign bit xor-ed with bit 31 of s.
LEV_HD float xor_sign(float x, uint32_t s) {
  uint32_t u;
  memcpy(&u, &x, sizeof u);
  u ^= s & 0x80000000u;
  memcpy(&x, &u, sizeof x);
  return x;
}

LEV_HD void sincos_bounded(float x, float* s, float* c) {
  constexpr float kTwoOverPi = 0x1.45f306p-1f;   // 2/pi rounded to float
  constexpr float kP1 = 0x1.921fb6p+0f;     // pi/2 rounded to float
  constexpr float kP2 = -0x1.777a5cp-25f;   // pi/2 - P1 rounded to float
  constexpr float kP3 = -0x1.ee59dap-50f;   // pi/2 - P1 - P2 rounded to float
  constexpr float S1 = -0x1.555546p-3f, S2 = 0x1.110732p-7f, S3 = -0x1.994078p-13f;
  constexpr float C1 = -0x1p-1f, C2 = 0x1.55553cp-5f, C3 = -0x1.6c07e2p-10f, C4 = 0x1.991314p-16f;
  const float k = rintf(x * kTwoOverPi);
  float r = fmaf(-k, kP1, x);
  r = fmaf(-k, kP2, r);
  r = fmaf(-k, kP3, r);
  const float r2 = r * r;
  float ps = fmaf(r2, S3, S2);
  ps = fmaf(r2, ps, S1);
  const float sr = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, C4, C3);
  pc = fmaf(r2, pc, C2);
  pc = fmaf(r2, pc, C1);
  const float cr = fmaf(r2, pc, 1.0f);
  const uint32_t q = (uint32_t)(int)k;
  const bool odd = (q & 1u) != 0;
  // quadrant q: sin x = (sin r, cos r, -sin r, -cos r), cos x = (cos r, -sin r, -cos r, sin r)
  *s = xor_sign(odd ? cr : sr, q << 30);
  *c = xor_sign(odd ? sr : cr, (q + 1u) << 30);
}

}  // namespace lev
