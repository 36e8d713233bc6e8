// step_kernel.h -- device helpers of the fused LEVIATHAN step (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "leviathan_internal.h"

#define LEV_HD __host__ __device__ __forceinline__
#include "sincos_bounded.h"

namespace lev {

// 2*pi rounded to FP32 (the reference's 2.0f * CUDART_PI_F, leviathan_kernel.cu:102).
constexpr float kTwoPi = 6.28318548f;
// uint16 exact-wrap angles: q in [0, 65536) stands for q * 2*pi / 65536 radians.
constexpr float kU16ToRad = 0x1.921fb6p-14f;      // 2*pi / 65536 rounded to FP32
constexpr float kU16ToRadTail = -0x1.777a5cp-39f; // 2*pi / 65536 - kU16ToRad, rounded to FP32
constexpr float kRadToU16 = 10430.3784f;      // 65536 / (2*pi)

// Exactly the reference's per-node wraps (true division): prep's [-pi, pi)
// (:105) and integrate's [0, 2*pi) (:188).
__device__ __forceinline__ float wrap_pm_pi_div(float x) {
  return fmaf(-floorf(x / kTwoPi + 0.5f), kTwoPi, x);
}
__device__ __forceinline__ float wrap_0_2pi_div(float x) {
  return fmaf(-floorf(x / kTwoPi), kTwoPi, x);
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
  // The angle q * 2*pi/65536 rounded once to FP32: q times the constant split
  // into an FP32 head (kU16ToRad) and tail, added by one fused multiply-add
  // (correctly rounded for all 65,536 codes).  The product with the head alone
  // is 2.8e-8 relative too large on average (the head's rounding), a bias that
  // adds up over a hub's ten thousand incoming edges.
  __device__ __forceinline__ static float decode(uint16_t q) {
    const float qf = (float)q;
    return fmaf(qf, kU16ToRad, qf * kU16ToRadTail);
  }
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

// One DPP move of a double (both halves) within rows of 16 lanes.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// FP64 sum over the 64 lanes of a wavefront (all lanes active), in a fixed
// order, returned to every lane.  Data moves by DPP within each row of 16
// lanes (lane xor 1, xor 2, then rotations by 4 and 8: lane 0 of a row ends
// with its row's sum) and by readlane across the four rows, with no LDS
// round trips.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_d<0x124>(v);  // row_ror:4
  v += dpp_d<0x128>(v);  // row_ror:8
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
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

// sin and cos of any float: the bounded routine, and the general one for
// |x| beyond its range and NaN (a branch no wave takes while every phase is
// wrapped).
__device__ __forceinline__ void sincos_fast(float x, float* s, float* c) {
  sincos_bounded(x, s, c);
  if (!(fabsf(x) <= kSinCosBoundedMax)) sincosf(x, s, c);
}

// Per-row part of the step, shared by every kernel variant so the arithmetic
// is written once.  The library is compiled with -ffp-contract=off and every
// fused multiply-add is explicit, so all kernel variants round identically.
// prep (leviathan_kernel.cu:101-109): eps from the wrapped prediction error;
// the predictor advances with the OLD omega_hat.
struct RowState {
  float th, eps, thh_new, omh_new, gate;
};

__device__ __forceinline__ RowState row_prep(float th, float thh, float omh, float dt, float kappa_w, float gamma) {
  RowState s;
  s.th = th;
  float unused;
  sincos_fast(wrap_pm_pi_div(th - thh), &s.eps, &unused);
  s.thh_new = fmaf(dt, omh, thh);
  s.omh_new = fmaf(kappa_w * dt, s.eps, omh);
  s.gate = fmaf(-gamma, fabsf(s.eps), 1.0f);
  return s;
}

// integrate (leviathan_kernel.cu:183-190): wrap to [0, 2*pi), NaN -> 0.
__device__ __forceinline__ float row_integrate(const RowState& s, float om, float g, float coupling, float dt,
                                               float alpha) {
  float x = fmaf(dt, fmaf(-alpha, s.eps, fmaf(g, coupling, om)), s.th);
  x = wrap_0_2pi_div(x);
  if (!__builtin_isfinite(x)) x = 0.0f;
  return x;
}

// The reference's guard and clamp of a new weight (leviathan_kernel.cu:159-163):
// NaN and infinities to w_min, then clamp to [w_min, w_max] (w_min <= w_max is
// checked when a handle is created).  The clamp is one v_med3_f32 of a finite
// value (fminf and fmaxf made the compiler quiet both bounds for every edge,
// and compares with selects took four instructions).
__device__ __forceinline__ float clamp_weight(float wv, float w_min, float w_max) {
  if (!__builtin_isfinite(wv)) wv = w_min;
  return __builtin_amdgcn_fmed3f(wv, w_min, w_max);
}

// Hebbian update of one edge (leviathan_kernel.cu:157-163); returns the new
// weight and writes the coupling contribution w_old * sin(delta).  The
// reference wraps delta to [-pi, pi) first (:146); sin and cos are
// 2*pi-periodic and, with both phases in [0, 2*pi), |delta| < 2*pi, which the
// bounded routine reduces directly (with pi/2 to 2^-76 rather than a float
// 2*pi), so the wrap is left out.
__device__ __forceinline__ float edge_update(float h, float th_i, float gate, float w, float dt, float eta, float lam,
                                             float w_min, float w_max, float* contrib) {
  const float diff = h - th_i;
  float sn, cs;
  sincos_fast(diff, &sn, &cs);
  *contrib = w * sn;
  float wv = fmaf(dt, fmaf(eta * cs, gate, -lam * w), w);
  return clamp_weight(wv, w_min, w_max);
}

// uint16 history: the delayed phase is exactly q / 65536 revolutions, so the
// edge takes sin and cos of delta = h - theta_i in revolutions (th_rev is
// theta_i / (2*pi), formed once per row) with the hardware v_sin_f32 and
// v_cos_f32, which take revolutions: three instructions where the bounded
// polynomial routine takes about thirty.  Measured at the 10^7-node graph:
// one step within 2.2e-7 (theta) and 2.8e-7 (W) of the FP64 oracle, and
// 2,000 steps' windowed r within 3.9e-3 of the FP64 run (3.7e-3 with the
// polynomial routine, whose error the uint16 storage error dwarfs).
// q is the uint16 code as a float (exact).
__device__ __forceinline__ float edge_update_rev(float q, float th_rev, float gate, float w, float dt, float eta,
                                                 float lam, float w_min, float w_max, float* contrib) {
  const float d = fmaf(q, 0x1p-16f, -th_rev);
  const float sn = __builtin_amdgcn_sinf(d);
  const float cs = __builtin_amdgcn_cosf(d);
  *contrib = w * sn;
  const float wv = fmaf(dt, fmaf(eta * cs, gate, -lam * w), w);
  return clamp_weight(wv, w_min, w_max);
}
constexpr float kInvTwoPi = 0.159154943f;  // 1 / (2*pi)

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
