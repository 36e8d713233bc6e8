// phi.hip -- hub transition statistics and the bipartition measure phi: the
// C-ABI of include/leviathan_phi.h.
//
// Reference: src/cuda/leviathan_phi_kernel.cu (kernels :60-235, host API
// :240-369).  Layout and kernels here:
//
//   counts      int64 [S * S], row-major by previous state; row totals int64 [S]
//   sbn         FP64 [H][S]: hub-major, so lanes of a wavefront read
//               consecutive states
//   partitions  FP64 [P], P = 2^(H-1) - 1, bipartition p has mask p + 1
//
//   phi_accumulate_kernel  one wavefront: hub_load + hub_record (phi_internal.h);
//                          the same code runs inside the step's finalize
//                          kernel when the handle is attached to an engine
//   phi_sbn_kernel         one workgroup per state s: integer sums of the
//                          counts over the targets with bit n set, then one
//                          FP64 division per hub (the exactly rounded
//                          conditional probability; the reference sums FP32
//                          quotients, :98-136)
//   phi_partition_kernel   one workgroup per bipartition, lanes over states,
//                          FP64, summed in a fixed order (:143-207)
//   phi_mip_kernel         one wavefront: minimum and its first mask (:213-235)
//
// Sizes are at most S = 1024: the whole table is 8 MiB and phi_compute is a
// few microseconds of work, so these kernels favour exactness (integer
// sums, FP64) over speed.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "leviathan_internal.h"
#include "leviathan_phi.h"
#include "step_kernel.h"

namespace lev {

constexpr int kPhiBlock = 256;

struct PhiResult {
  double phi;
  int32_t mask;
  int32_t pad;
};

__global__ __launch_bounds__(64) void phi_accumulate_kernel(PhiDev p, const float* theta) {
  const HubLoad h = hub_load(p, theta, threadIdx.x);
  hub_record(p, h, threadIdx.x);
}

__global__ __launch_bounds__(kPhiBlock) void phi_sbn_kernel(const int64_t* counts, const int64_t* row_totals,
                                                            int32_t H, double* sbn) {
  __shared__ long long s_red[kPhiMaxHubs][kPhiBlock / 64];
  const int32_t S = 1 << H;
  const int32_t s = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  long long acc[kPhiMaxHubs];
#pragma unroll
  for (int n = 0; n < kPhiMaxHubs; ++n) acc[n] = 0;
  for (int32_t t = tid; t < S; t += kPhiBlock) {
    const long long c = counts[(int64_t)s * S + t];
#pragma unroll
    for (int n = 0; n < kPhiMaxHubs; ++n) acc[n] += ((t >> n) & 1) ? c : 0;
  }
#pragma unroll
  for (int n = 0; n < kPhiMaxHubs; ++n) {
    long long v = acc[n];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) s_red[n][wave] = v;
  }
  __syncthreads();
  if (tid < H) {
    long long num = 0;
    for (int w = 0; w < kPhiBlock / 64; ++w) num += s_red[tid][w];
    const long long tot = row_totals[s];
    // a state never left: the uniform row 1/S summed over the S/2 targets with the bit set
    sbn[(int64_t)tid * S + s] = tot > 0 ? (double)num / (double)tot : 0.5;
  }
}

__global__ __launch_bounds__(kPhiBlock) void phi_partition_kernel(const double* sbn, int32_t H, double* part) {
  __shared__ double s_red[kPhiBlock / 64];
  const int32_t S = 1 << H;
  const int32_t mask = blockIdx.x + 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nA = __popc((unsigned)(mask & (S - 1)));
  if (nA == 0 || nA == H) {  // not a bipartition (never the case for mask < 2^(H-1))
    if (tid == 0) part[blockIdx.x] = 1e30;
    return;
  }
  double acc = 0.0;
  for (int32_t s = tid; s < S; s += kPhiBlock) {
    double pi = 1.0, pa = 1.0, pb = 1.0;
    for (int n = 0; n < H; ++n) {
      const double pn = sbn[(int64_t)n * S + s];
      double prob = ((s >> n) & 1) ? pn : 1.0 - pn;
      prob = fmax(prob, 1e-10);
      pi *= prob;
      if ((mask >> n) & 1) {
        pa *= prob;
      } else {
        pb *= prob;
      }
    }
    const double pp = pa * pb;
    acc += fabs(log(fmax(pi, 1e-30)) - log(fmax(pp, 1e-30)));
  }
  acc = wave_sum(acc);
  if (lane == 0) s_red[wave] = acc;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int w = 0; w < kPhiBlock / 64; ++w) t += s_red[w];
    part[blockIdx.x] = t / (double)S;  // S is a power of two: exact
  }
}

__global__ __launch_bounds__(64) void phi_mip_kernel(const double* part, int32_t P, PhiResult* out) {
  const int lane = threadIdx.x;
  double best = 1e30;
  int32_t idx = -1;
  for (int32_t p = lane; p < P; p += 64)
    if (idx < 0 || part[p] < best) {  // strict: a lane keeps its first minimum
      best = part[p];
      idx = p;
    }
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_xor(best, off);
    const int32_t oi = __shfl_xor(idx, off);
    // smaller value wins; ties (and lanes with nothing) go to the lower index
    const bool take = oi >= 0 && (idx < 0 || ob < best || (ob == best && oi < idx));
    if (take) {
      best = ob;
      idx = oi;
    }
  }
  if (lane == 0) {
    PhiResult r;
    r.phi = idx < 0 ? 1e30 : best;
    r.mask = idx < 0 ? 0 : idx + 1;
    r.pad = 0;
    *out = r;
  }
}

__global__ void phi_set_state_kernel(PhiState* st, int32_t state, int64_t calls) {
  PhiState v;
  v.prev = state;
  v.last = state;
  v.calls = calls;
  *st = v;
}

}  // namespace lev

struct PhiHandle {
  int device = 0;
  int32_t H = 0, S = 0, P = 0;
  int32_t max_hub = 0;
  lev::PhiState* d_st = nullptr;
  int64_t* d_counts = nullptr;
  int64_t* d_row_totals = nullptr;
  double* d_sbn = nullptr;
  double* d_part = nullptr;
  lev::PhiResult* d_result = nullptr;
  float* d_hub_theta = nullptr;
  void* h_pinned = nullptr;  // result and staging for the host phases
  std::vector<int32_t> hubs;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  LeviathanHandle* engine = nullptr;  // attached engine, if any
  hipStream_t engine_stream = nullptr;
};

namespace lev {

PhiDev phi_view(const PhiHandle* p) {
  PhiDev v{};
  for (int i = 0; i < p->H; ++i) v.hub[i] = p->hubs[i];
  v.st = p->d_st;
  v.counts = p->d_counts;
  v.row_totals = p->d_row_totals;
  v.num_hubs = p->H;
  return v;
}

int32_t phi_max_hub(const PhiHandle* p) { return p->max_hub; }
int phi_device(const PhiHandle* p) { return p->device; }
LeviathanHandle* phi_engine(const PhiHandle* p) { return p->engine; }
void phi_link(PhiHandle* p, LeviathanHandle* engine, hipStream_t engine_stream) {
  p->engine = engine;
  p->engine_stream = engine_stream;
}

}  // namespace lev

namespace {

void phi_err(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  lev::error_set(buf);
}

#define PHI_HIP(call, fail)                                                                    \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      phi_err("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__);      \
      return fail;                                                                             \
    }                                                                                          \
  } while (0)

bool check(PhiHandle* h) {
  lev::error_clear();
  if (!h) {
    phi_err("null PhiHandle");
    return false;
  }
  if (hipSetDevice(h->device) != hipSuccess) {
    phi_err("hipSetDevice(%d) failed", h->device);
    return false;
  }
  return true;
}

// Order the handle's stream after the attached engine's work.
bool after_engine(PhiHandle* h) {
  if (!h->engine_stream) return true;
  PHI_HIP(hipEventRecord(h->ev, h->engine_stream), false);
  PHI_HIP(hipStreamWaitEvent(h->stream, h->ev, 0), false);
  return true;
}

void destroy(PhiHandle* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void* ptrs[] = {h->d_st, h->d_counts, h->d_row_totals, h->d_sbn, h->d_part, h->d_result,
                  h->d_hub_theta};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (h->h_pinned) (void)hipHostFree(h->h_pinned);
  if (h->ev) (void)hipEventDestroy(h->ev);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

template <class T>
bool alloc(T** p, size_t count) {
  PHI_HIP(hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1)), false);
  return true;
}

bool zero_state(PhiHandle* h) {
  PHI_HIP(hipMemsetAsync(h->d_counts, 0, sizeof(int64_t) * (size_t)h->S * h->S, h->stream), false);
  PHI_HIP(hipMemsetAsync(h->d_row_totals, 0, sizeof(int64_t) * (size_t)h->S, h->stream), false);
  lev::phi_set_state_kernel<<<1, 1, 0, h->stream>>>(h->d_st, -1, 0);
  PHI_HIP(hipGetLastError(), false);
  return true;
}

// sbn, partitions and the minimum into d_result, then the result to pinned memory.
bool run_compute(PhiHandle* h) {
  if (!after_engine(h)) return false;
  lev::phi_sbn_kernel<<<h->S, lev::kPhiBlock, 0, h->stream>>>(h->d_counts, h->d_row_totals, h->H, h->d_sbn);
  PHI_HIP(hipGetLastError(), false);
  if (h->P > 0) {
    lev::phi_partition_kernel<<<h->P, lev::kPhiBlock, 0, h->stream>>>(h->d_sbn, h->H, h->d_part);
    PHI_HIP(hipGetLastError(), false);
  }
  lev::phi_mip_kernel<<<1, 64, 0, h->stream>>>(h->d_part, h->P, h->d_result);
  PHI_HIP(hipGetLastError(), false);
  PHI_HIP(hipMemcpyAsync(h->h_pinned, h->d_result, sizeof(lev::PhiResult), hipMemcpyDeviceToHost, h->stream), false);
  return true;
}

bool read_state(PhiHandle* h, lev::PhiState* out) {
  if (!after_engine(h)) return false;
  PHI_HIP(hipMemcpyAsync(h->h_pinned, h->d_st, sizeof(lev::PhiState), hipMemcpyDeviceToHost, h->stream), false);
  PHI_HIP(hipStreamSynchronize(h->stream), false);
  std::memcpy(out, h->h_pinned, sizeof(lev::PhiState));
  return true;
}

}  // namespace

extern "C" {

PhiHandle* phi_init(int num_hubs, int* h_hub_indices) {
  lev::error_clear();
  if (num_hubs < 1 || num_hubs > lev::kPhiMaxHubs) {
    phi_err("phi_init: num_hubs must be in [1, %d] (got %d)", lev::kPhiMaxHubs, num_hubs);
    return nullptr;
  }
  if (!h_hub_indices) {
    phi_err("phi_init: null hub index array");
    return nullptr;
  }
  PhiHandle* h = new PhiHandle();
  h->H = num_hubs;
  h->S = 1 << num_hubs;
  h->P = (1 << (num_hubs - 1)) - 1;
  h->hubs.assign(h_hub_indices, h_hub_indices + num_hubs);
  for (int i = 0; i < num_hubs; ++i) {
    if (h->hubs[i] < 0) {
      phi_err("phi_init: hub index %d is negative (%d)", i, h->hubs[i]);
      delete h;
      return nullptr;
    }
    h->max_hub = std::max(h->max_hub, h->hubs[i]);
  }
  auto fail = [&]() -> PhiHandle* {
    destroy(h);
    return nullptr;
  };
  if (hipGetDevice(&h->device) != hipSuccess) {
    phi_err("phi_init: hipGetDevice failed");
    return fail();
  }
  if (!alloc(&h->d_st, 1) || !alloc(&h->d_counts, (size_t)h->S * h->S) ||
      !alloc(&h->d_row_totals, h->S) || !alloc(&h->d_sbn, (size_t)h->S * h->H) || !alloc(&h->d_part, h->P) ||
      !alloc(&h->d_result, 1) || !alloc(&h->d_hub_theta, lev::kPhiMaxHubs))
    return fail();
  if (hipHostMalloc(&h->h_pinned, 4096, hipHostMallocDefault) != hipSuccess) {
    h->h_pinned = nullptr;
    phi_err("phi_init: hipHostMalloc failed");
    return fail();
  }
  if (hipStreamCreate(&h->stream) != hipSuccess || hipEventCreateWithFlags(&h->ev, hipEventDisableTiming) != hipSuccess) {
    phi_err("phi_init: stream or event creation failed");
    return fail();
  }
  if (!zero_state(h) || hipStreamSynchronize(h->stream) != hipSuccess) {
    if (lev::error_empty()) phi_err("phi_init: initial reset failed");
    return fail();
  }
  return h;
}

void phi_accumulate(PhiHandle* h, const float* d_theta) {
  if (!check(h)) return;
  if (h->engine) {
    phi_err("phi_accumulate: the handle is attached to an engine, whose steps record the states");
    return;
  }
  if (!d_theta) {
    phi_err("phi_accumulate: null phase array");
    return;
  }
  // asynchronous on the handle's stream, like the reference's (:290-299); later calls wait for it
  lev::phi_accumulate_kernel<<<1, 64, 0, h->stream>>>(lev::phi_view(h), d_theta);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) phi_err("phi_accumulate: %s", hipGetErrorString(e));
}

int phi_accumulate_host(PhiHandle* h, const float* h_theta, int64_t n) {
  if (!check(h)) return -1;
  if (h->engine) {
    phi_err("phi_accumulate_host: the handle is attached to an engine, whose steps record the states");
    return -1;
  }
  if (!h_theta || n <= h->max_hub) {
    phi_err("phi_accumulate_host: %lld phases do not cover hub index %d", (long long)n, h->max_hub);
    return -1;
  }
  PHI_HIP(hipStreamSynchronize(h->stream), -1);  // the staging buffer may still be in use
  float* stage = static_cast<float*>(h->h_pinned) + 64;
  for (int i = 0; i < h->H; ++i) stage[i] = h_theta[h->hubs[i]];
  PHI_HIP(hipMemcpyAsync(h->d_hub_theta, stage, sizeof(float) * h->H, hipMemcpyHostToDevice, h->stream), -1);
  lev::PhiDev v = lev::phi_view(h);
  for (int i = 0; i < h->H; ++i) v.hub[i] = i;  // the staged phases are in hub order
  lev::phi_accumulate_kernel<<<1, 64, 0, h->stream>>>(v, h->d_hub_theta);
  PHI_HIP(hipGetLastError(), -1);
  PHI_HIP(hipStreamSynchronize(h->stream), -1);
  return 0;
}

float phi_compute(PhiHandle* h) {
  double phi = 0.0;
  if (phi_compute_ex(h, &phi, nullptr, nullptr, nullptr) != 0) return std::numeric_limits<float>::quiet_NaN();
  return (float)phi;
}

int phi_compute_ex(PhiHandle* h, double* h_phi, int* h_mip_mask, double* h_sbn_tpm, double* h_partitions) {
  if (!check(h)) return -1;
  if (!h_phi) {
    phi_err("phi_compute_ex: null output");
    return -1;
  }
  if (!run_compute(h)) return -1;
  std::vector<double> sbn;
  if (h_sbn_tpm) {
    sbn.resize((size_t)h->S * h->H);
    PHI_HIP(hipMemcpyAsync(sbn.data(), h->d_sbn, sizeof(double) * sbn.size(), hipMemcpyDeviceToHost, h->stream), -1);
  }
  if (h_partitions && h->P > 0)
    PHI_HIP(hipMemcpyAsync(h_partitions, h->d_part, sizeof(double) * h->P, hipMemcpyDeviceToHost, h->stream), -1);
  PHI_HIP(hipStreamSynchronize(h->stream), -1);
  lev::PhiResult r;
  std::memcpy(&r, h->h_pinned, sizeof r);
  *h_phi = r.phi;
  if (h_mip_mask) *h_mip_mask = r.mask;
  if (h_sbn_tpm)  // device layout is hub-major: [n][s] -> [s][n]
    for (int32_t s = 0; s < h->S; ++s)
      for (int32_t n = 0; n < h->H; ++n) h_sbn_tpm[(size_t)s * h->H + n] = sbn[(size_t)n * h->S + s];
  return 0;
}

void phi_reset(PhiHandle* h) {
  if (!check(h)) return;
  if (!after_engine(h)) return;
  if (!zero_state(h)) return;
  const hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) phi_err("phi_reset: %s", hipGetErrorString(e));
}

int64_t phi_get_transition_count(PhiHandle* h) {
  if (!check(h)) return -1;
  lev::PhiState st;
  if (!read_state(h, &st)) return -1;
  return st.calls;
}

int phi_get_state(PhiHandle* h, int* h_state, int64_t* h_calls) {
  if (!check(h)) return -1;
  lev::PhiState st;
  if (!read_state(h, &st)) return -1;
  if (h_state) *h_state = st.last;
  if (h_calls) *h_calls = st.calls;
  return 0;
}

int phi_get_counts(PhiHandle* h, int64_t* h_counts, int64_t* h_row_totals) {
  if (!check(h)) return -1;
  if (!after_engine(h)) return -1;
  if (h_counts)
    PHI_HIP(hipMemcpyAsync(h_counts, h->d_counts, sizeof(int64_t) * (size_t)h->S * h->S, hipMemcpyDeviceToHost,
                           h->stream),
            -1);
  if (h_row_totals)
    PHI_HIP(hipMemcpyAsync(h_row_totals, h->d_row_totals, sizeof(int64_t) * (size_t)h->S, hipMemcpyDeviceToHost,
                           h->stream),
            -1);
  PHI_HIP(hipStreamSynchronize(h->stream), -1);
  return 0;
}

int phi_set_state(PhiHandle* h, const int64_t* h_counts, const int64_t* h_row_totals, int state, int64_t calls) {
  if (!check(h)) return -1;
  if (!h_counts || !h_row_totals) {
    phi_err("phi_set_state: null counts");
    return -1;
  }
  if (state < -1 || state >= h->S || calls < 0) {
    phi_err("phi_set_state: state %d or call count %lld out of range", state, (long long)calls);
    return -1;
  }
  for (int32_t s = 0; s < h->S; ++s) {
    int64_t sum = 0;
    for (int32_t t = 0; t < h->S; ++t) {
      const int64_t c = h_counts[(size_t)s * h->S + t];
      if (c < 0) {
        phi_err("phi_set_state: negative count at (%d, %d)", s, t);
        return -1;
      }
      sum += c;
    }
    if (sum != h_row_totals[s]) {
      phi_err("phi_set_state: row %d sums to %lld, its total is %lld", s, (long long)sum,
              (long long)h_row_totals[s]);
      return -1;
    }
  }
  if (!after_engine(h)) return -1;
  PHI_HIP(hipMemcpyAsync(h->d_counts, h_counts, sizeof(int64_t) * (size_t)h->S * h->S, hipMemcpyHostToDevice,
                         h->stream),
          -1);
  PHI_HIP(hipMemcpyAsync(h->d_row_totals, h_row_totals, sizeof(int64_t) * (size_t)h->S, hipMemcpyHostToDevice,
                         h->stream),
          -1);
  lev::phi_set_state_kernel<<<1, 1, 0, h->stream>>>(h->d_st, state, calls);
  PHI_HIP(hipGetLastError(), -1);
  PHI_HIP(hipStreamSynchronize(h->stream), -1);
  return 0;
}

void phi_free(PhiHandle* h) {
  lev::error_clear();
  if (!h) return;
  if (h->engine) (void)leviathan_attach_phi(h->engine, nullptr);
  lev::error_clear();
  destroy(h);
}

}  // extern "C"
