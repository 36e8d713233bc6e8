// phi_internal.h -- device state of the hub transition statistics (the
// PhiHandle of include/leviathan_phi.h), shared by phi.hip (its own kernels
// and host API) and leviathan_kernels.hip (the finalize kernel, which records
// a transition after every step of an engine the handle is attached to).
//
// Reference: src/cuda/leviathan_phi_kernel.cu:27-53 (PhiHandle) and :60-92
// (binarize_and_accumulate_kernel).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct PhiHandle;
struct LeviathanHandle;

namespace lev {

constexpr int kPhiMaxHubs = 10;  // leviathan_phi_kernel.cu:243-246: S = 2^H <= 1024 states

struct PhiState {
  int32_t prev;   // state recorded by the last call, -1 before the first
  int32_t last;   // the same state, as last returned (the reference's d_state_out)
  int64_t calls;  // accumulate calls since the last reset (the reference's transition_counter)
};

struct PhiDev {
  PhiState* st;         // null: no handle attached
  int64_t* counts;      // [S * S] transitions prev -> state, row-major by prev
  int64_t* row_totals;  // [S]
  int32_t num_hubs;
  int32_t hub[kPhiMaxHubs];  // node index of hub i (carried in the kernel arguments)
};

// Hub state of one phase vector and its transition from the previous state,
// for one full wavefront in two parts, so that a kernel can issue the loads
// early and hide their latency behind other work: lane i < H reads hub i's
// phase, lane 0 the previous state.
struct HubLoad {
  float th;
  PhiState st;
};

__device__ __forceinline__ HubLoad hub_load(const PhiDev& p, const float* theta, int lane) {
  HubLoad h;
  h.th = 0.0f;
  h.st = PhiState{-1, -1, 0};
  int32_t node = 0;
#pragma unroll
  for (int i = 0; i < kPhiMaxHubs; ++i)
    if (lane == i) node = p.hub[i];
  if (lane < p.num_hubs) h.th = theta[node];
  if (lane == 0) h.st = *p.st;
  return h;
}

// Bit i is set when hub i's phase, reduced by fmodf into [0, 2*pi), exceeds
// pi, with the reference's FP32 constants (leviathan_phi_kernel.cu:69-79:
// 6.2831853f rounds to the float 6.28318548 and 3.1415926f to 3.14159250;
// fmodf is exact).  Lane 0 adds the transition (atomics without return, so
// the wavefront does not wait for them) and stores the new state; calls on
// one stream are ordered, so the state's read-modify-write needs no atomic.
__device__ __forceinline__ void hub_record(const PhiDev& p, const HubLoad& h, int lane) {
  const float kTwoPi = 6.2831853f;
  const float kPi = 3.1415926f;
  bool bit = false;
  if (lane < p.num_hubs) {
    float x = fmodf(h.th, kTwoPi);
    if (x < 0.0f) x += kTwoPi;
    bit = x > kPi;
  }
  const uint64_t m = __ballot(bit);
  if (lane == 0) {
    const int32_t S = 1 << p.num_hubs;
    const int32_t state = (int32_t)(m & (uint64_t)(S - 1));
    const int32_t prev = h.st.prev;
    if (prev >= 0 && prev < S) {
      atomicAdd(reinterpret_cast<unsigned long long*>(p.counts + (int64_t)prev * S + state), 1ull);
      atomicAdd(reinterpret_cast<unsigned long long*>(p.row_totals + prev), 1ull);
    }
    *p.st = PhiState{state, state, h.st.calls + 1};
  }
}

// Host-side accessors of a PhiHandle for the engine (defined in phi.hip).
PhiDev phi_view(const PhiHandle* p);
int32_t phi_max_hub(const PhiHandle* p);
int phi_device(const PhiHandle* p);
LeviathanHandle* phi_engine(const PhiHandle* p);
void phi_link(PhiHandle* p, LeviathanHandle* engine, hipStream_t engine_stream);

// The library's per-thread error message (leviathan_last_error), shared by
// every source file of the C-ABI.
void error_set(const char* msg);
void error_clear();
bool error_empty();

}  // namespace lev
