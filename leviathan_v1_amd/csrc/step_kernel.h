// step_kernel.h -- device helpers of the fused LEVIATHAN step (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "leviathan_internal.h"

namespace lev {

// 2*pi rounded to FP32 (the reference's 2.0f * CUDART_PI_F, leviathan_kernel.cu:102).
constexpr float kTwoPi = 6.28318548f;
constexpr float kInvTwoPi = 0.159154943f;
// uint16 exact-wrap angles: q in [0, 65536) stands for q * 2*pi / 65536 radians.
constexpr float kU16ToRad = 9.58737992e-05f;  // 2*pi / 65536
constexpr float kRadToU16 = 10430.3784f;      // 65536 / (2*pi)

// Wrap to [-pi, pi) for the coupling argument.  The reference divides by 2*pi
// (:146); a multiply by the rounded reciprocal changes only which side of the
// +-pi seam a value lands on, which sin/cos do not see.
__device__ __forceinline__ float wrap_pm_pi(float x) {
  return x - floorf(x * kInvTwoPi + 0.5f) * kTwoPi;
}

// Exactly the reference's per-node wraps (true division): prep's [-pi, pi)
// (:105) and integrate's [0, 2*pi) (:188).
__device__ __forceinline__ float wrap_pm_pi_div(float x) {
  return x - floorf(x / kTwoPi + 0.5f) * kTwoPi;
}
__device__ __forceinline__ float wrap_0_2pi_div(float x) {
  return x - floorf(x / kTwoPi) * kTwoPi;
}

template <class H>
struct HistCodec;

template <>
struct HistCodec<float> {
  __device__ __forceinline__ static float decode(float v) { return v; }
  __device__ __forceinline__ static float encode(float x) { return x; }
};

template <>
struct HistCodec<uint16_t> {
  __device__ __forceinline__ static float decode(uint16_t q) { return (float)q * kU16ToRad; }
  // Round to the nearest 2*pi/65536 step and wrap modulo 65536 (so 2*pi -> 0).
  __device__ __forceinline__ static uint16_t encode(float x) {
    return (uint16_t)((uint32_t)__float2int_rn(x * kRadToU16) & 0xFFFFu);
  }
};

template <int LAYOUT>
__device__ __forceinline__ int64_t ring_index(int64_t pitch, int32_t node, int32_t slot) {
  if constexpr (LAYOUT == 0) {
    return (int64_t)node * pitch + slot;  // node-major: one node's history is contiguous
  } else {
    return (int64_t)slot * pitch + node;  // slot-major: one time slot is contiguous
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Neumaier's compensated summation (Neumaier 1974), one term at a time.
struct CompSum {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float v) {
    const float t = s + v;
    if (fabsf(s) >= fabsf(v)) {
      c += (s - t) + v;
    } else {
      c += (v - t) + s;
    }
    s = t;
  }
  __device__ __forceinline__ float value() const { return s + c; }
};

// Logical row-block index for a workgroup.  Workgroups are dealt to the 8
// XCDs round-robin; this mapping gives each XCD one contiguous run of row
// blocks so neighbouring rows (which share source nodes) meet in one L2.
__device__ __forceinline__ int32_t xcd_block(int32_t bid, int32_t nb, int32_t remap) {
  if (!remap) return bid;
  const int32_t xcd = bid % kNumXcd;
  const int32_t local = bid / kNumXcd;
  const int32_t base = nb / kNumXcd;
  const int32_t rem = nb % kNumXcd;
  return xcd * base + (xcd < rem ? xcd : rem) + local;
}

}  // namespace lev
