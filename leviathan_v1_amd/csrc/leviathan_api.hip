// leviathan_api.hip -- host side of the C-ABI declared in include/leviathan.h.
//
// Reference counterpart: the host API of src/cuda/leviathan_kernel.cu:266-497
// (LeviathanHandle, leviathan_init/step/get_theta/set_theta/free/set_gain/
// get_gain/set_gain_controller/get_vram_usage/reset_weights).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "leviathan.h"
#include "leviathan_internal.h"
#include "leviathan_phi.h"
#include "plan.h"

using lev::FinalizeArgs;
using lev::NetCtrl;
using lev::StepArgs;

namespace {

thread_local std::string g_last_error;

void set_err(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

#define LEV_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      set_err("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return false;                                                                        \
    }                                                                                      \
  } while (0)

// Device scratch of leviathan_features_colsum_noalloc: the staged row weights
// (kNoAllocRows doubles) followed by the output (kNoAllocFeatures doubles).
// Each GPU holds its own copy until the process ends.  Host code never passes
// the array's name to a kernel or a copy (on the host it is only a stand-in):
// it uses the address hipGetSymbolAddress returns per device (scratch_ptr).
constexpr int64_t kNoAllocRows = 65536;
constexpr int32_t kNoAllocFeatures = 8192;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocRows + kNoAllocFeatures];

std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Looks up (once per device, with that device current) and returns the device
// address of this GPU's scratch array, or null after recording the error.
double* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) {
    set_err("device %d is beyond the %d devices the allocation-free scratch table holds", device, kMaxDevices);
    return nullptr;
  }
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_err("hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_dev[device] = static_cast<double*>(p);
  }
  return g_scratch_dev[device];
}

struct GraphEntry {
  uint32_t dt_bits;
  int chunk;
  bool capture;
  hipGraphExec_t exec;
};

// One schedule's work decomposition on the device.
// Plans 0-2: direct with 4, 8, 16 edges per lane; 3-5: binned and 6-8:
// staged, on the same row blocks.
constexpr int kPlanDirect4 = 0, kPlanDirect8 = 1, kPlanDirect16 = 2, kPlanBinned = 3, kPlanStaged = 6;
static_assert(LEV_NUM_PLANS == 9, "plan table");
constexpr int kNumPlans = LEV_NUM_PLANS;

struct DevPlan {
  bool built = false;
  int kind = LEV_SCHEDULE_DIRECT;
  int ept = 8;
  int4* d_items = nullptr;
  int32_t n_items = 0;
  int4* d_ranges = nullptr;
  int32_t xcd_split[lev::kNumXcd + 1] = {};
  bool owns = true;            // binned and staged plans share the row blocks of the direct plan with the same ept
  uint32_t* d_pos = nullptr;   // two-phase plans: [E] value slots in the scatter kernel's write order
  // two-phase plans, [E]: CSC edge -> index in its source block's slot order (low half) and,
  // per slot-order index, its run of consecutive slots (high half)
  uint32_t* d_rank_run = nullptr;
  int32_t* d_run_ptr = nullptr;
  int32_t* d_run_base = nullptr;
  uint16_t* d_lpos = nullptr;  // staged plans: [E] CSR edge -> slot within its block's range
  int4* d_rg_chunks = nullptr; // staged plans: regroup chunks, ranks and runs
  uint16_t* d_rg_rank = nullptr;
  int2* d_rg_runs = nullptr;
  int32_t n_rg_chunks = 0;
  // two-phase plans: each early edge's slot in the array the row pass reads (multi-step scatter)
  uint32_t* d_early_pos = nullptr;
  // staged plans: the early edges of each row block, gathered by the second pair of a block's
  // paired row passes (block_early_edges)
  int32_t* d_pe_ptr = nullptr;
  int2* d_pe = nullptr;
  bool owns_pos = false;
};

int plan_of_ept(int ept) { return ept == 4 ? kPlanDirect4 : ept == 16 ? kPlanDirect16 : kPlanDirect8; }
int plan_base(int kind) { return kind == LEV_SCHEDULE_BINNED ? kPlanBinned : kind == LEV_SCHEDULE_STAGED ? kPlanStaged : 0; }

}  // namespace

struct LeviathanHandle {
  LeviathanConfig cfg{};
  int device = 0;
  int32_t N = 0;
  int64_t E = 0;
  int32_t max_delay = 0;
  int32_t min_delay = 0;    // smallest edge delay (the paired row pass needs >= 1)
  int32_t n_slots = 0;      // S = max_delay + 2
  int64_t ring_pitch = 0;
  size_t ring_bytes = 0;    // the history ring's size
  int32_t n_nets = 1;
  int64_t step = 0;         // host mirror of the device step counter
  size_t vram = 0;

  int32_t* d_row_ptr = nullptr;
  int32_t* d_col = nullptr;
  uint8_t* d_delay = nullptr;
  float* d_w = nullptr;
  float* d_theta = nullptr;
  float* d_theta_hat = nullptr;
  float* d_omega = nullptr;
  float* d_omega_hat = nullptr;
  void* d_ring = nullptr;
  double2* d_partials = nullptr;
  int32_t n_partials = 0;
  double2* d_chunks = nullptr;  // chunk sums of the partials (finalize of one network)
  unsigned int* d_fin_done = nullptr;  // finalize of several networks: workgroups done this step
  DevPlan plans[kNumPlans];
  int active = kPlanDirect8;
  // binned schedule arrays
  int32_t* d_csc_ptr = nullptr;
  uint8_t* d_delay_csc = nullptr;
  uint16_t* d_dsrc = nullptr;      // slot-major ring: each CSC edge's source index in its block << 8 | delay
  uint32_t* d_pos = nullptr;
  void* d_hbuf = nullptr;    // two-phase schedules: the scatter kernel's bucketed values
  void* d_hbuf2 = nullptr;   // staged schedule: the values regrouped by row block
  uint32_t* d_hpos = nullptr;
  uint32_t* d_rank_run = nullptr;  // binned plans' ranks and scatter runs
  int32_t* d_run_ptr = nullptr;
  int32_t* d_run_base = nullptr;
  // multi-step scatter: the edges with delay < kMaxScatterSteps - 1 by (delay, source)
  int32_t* d_early_src = nullptr;
  uint8_t* d_early_delay = nullptr;
  uint32_t* d_early_pos = nullptr;  // binned plans' value slots of those edges
  int32_t early_cnt[lev::kMaxScatterSteps] = {};  // early_cnt[k] = edges with delay < k
  int scatter_steps = 1;   // consecutive steps one scatter pass gathers (two-phase schedules)
  int row_steps = 1;       // consecutive steps one row pass covers: 2 = paired row passes (leviathan_set_row_steps)
  float2* d_pair = nullptr;  // paired row pass: per node the second step's coupling sum and prediction error
  int hbuf_steps = 1;      // steps d_hbuf (and the staged schedule's d_hbuf2) have room for
  int4* d_sources = nullptr;  // phase-1 work items
  int32_t n_sources = 0;
  int32_t n_buckets = 0;
  NetCtrl* d_ctrl = nullptr;
  int64_t* d_step = nullptr;
  float* d_r_hist = nullptr;
  float* d_scratch = nullptr;   // N floats (history export)
  float* d_drive_b = nullptr;
  float* d_drive_u = nullptr;
  bool drive_on = false;

  // reservoir features
  int32_t* d_feat_nodes = nullptr;
  int32_t feat_n = 0;
  int64_t feat_max_rows = 0;
  int64_t feat_rows = 0;
  float* d_X = nullptr;
  int64_t* d_feat_base = nullptr;
  double* d_cs_partial = nullptr;  // kColsumChunks * 2 * feat_n
  // readout: FP64 Cholesky factor of the last leviathan_features_factor (rocSOLVER)
  rocblas_handle blas = nullptr;
  double* d_chol = nullptr;       // [chol_n * chol_n], lower triangle
  int32_t chol_n = 0;
  double* d_vec = nullptr;        // [2 * chol_n] right-hand side, solution
  bool chol_inverse = false;      // d_chol holds A^{-1} (matrix-core Grams), else the factor (FP64 Gram)
  double* d_dmean = nullptr;      // mean of the Gram's diagonal
  rocblas_int* d_info = nullptr;
  bool capturing = false;

  float* h_pinned = nullptr;    // kRHistSteps * n_nets floats
  hipStream_t stream = nullptr;
  std::vector<NetCtrl> h_ctrl;  // host copy of the static controller fields
  std::vector<GraphEntry> graphs;
  PhiHandle* phi = nullptr;     // hub statistics recorded by every step (leviathan_attach_phi)
};

namespace {

template <class T>
bool dalloc(LeviathanHandle* h, T** p, size_t count) {
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc((void**)p, bytes);
  // an empty array still gets one zeroed element: the row pass clamps its
  // loads into each block and reads element 0 of an empty edge array
  if (e == hipSuccess && count == 0) e = hipMemset(*p, 0, bytes);
  if (e != hipSuccess) {
    set_err("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    *p = nullptr;
    return false;
  }
  h->vram += bytes;
  return true;
}

// The scatter's per-edge rank and per-slot-order-index run, packed rank | run << 16 (both are
// indexed by the block's first out-edge + the lane's index), and its run tables.
// With the scatter's window form (hsc non-null) the per-edge words are its
// codes (window_scatter_codes) instead of rank | run.
bool upload_runs(LeviathanHandle* h, const std::vector<uint16_t>& rank, const lev::HostScatterRuns& r,
                 uint32_t** rank_run, int32_t** ptr, int32_t** base, const lev::HostScatter* hsc = nullptr) {
  if (rank.size() != r.run_idx.size()) {
    set_err("scatter ranks and runs differ in length");
    return false;
  }
  if (!dalloc(h, rank_run, rank.size()) || !dalloc(h, ptr, r.run_ptr.size()) || !dalloc(h, base, r.run_base.size()))
    return false;
  std::vector<uint32_t> packed(rank.size());
  if (hsc)
    lev::window_scatter_codes(*hsc, rank, r, packed);
  else
    for (size_t c = 0; c < rank.size(); ++c) packed[c] = (uint32_t)rank[c] | (uint32_t)r.run_idx[c] << 16;
  hipError_t e = hipSuccess;
  if (!packed.empty())
    e = hipMemcpy(*rank_run, packed.data(), sizeof(uint32_t) * packed.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(*ptr, r.run_ptr.data(), sizeof(int32_t) * r.run_ptr.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(*base, r.run_base.data(), sizeof(int32_t) * r.run_base.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_err("copy scatter runs: %s", hipGetErrorString(e));
    return false;
  }
  return true;
}

// Destroys the captured graphs.  A graph launched earlier may still be running
// on the handle's stream, so the stream is drained first: destroying an
// executable graph in flight would free the kernel arguments it still reads.
void drop_graphs(LeviathanHandle* h) {
  if (h->graphs.empty()) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
  h->graphs.clear();
}

void destroy(LeviathanHandle* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->phi) lev::phi_link(h->phi, nullptr, nullptr);
  drop_graphs(h);
  void* ptrs[] = {h->d_row_ptr, h->d_col,       h->d_delay,   h->d_w,        h->d_theta,   h->d_theta_hat,
                  h->d_omega,   h->d_omega_hat, h->d_ring,    h->d_partials, h->d_chunks, h->d_fin_done, h->d_ctrl,    h->d_step,
                  h->d_r_hist,  h->d_scratch,   h->d_drive_b, h->d_drive_u,  h->d_csc_ptr, h->d_delay_csc, h->d_dsrc,
                  h->d_pos,     h->d_hbuf,      h->d_hbuf2,   h->d_hpos,    h->d_rank_run, h->d_run_ptr, h->d_run_base, h->d_sources,  h->d_feat_nodes,
                  h->d_early_src, h->d_early_delay, h->d_early_pos, h->d_pair,
                  h->d_X,       h->d_feat_base, h->d_cs_partial};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& pl : h->plans) {
    if (pl.owns_pos) {
      void* pp[] = {pl.d_pos,   pl.d_rank_run,  pl.d_run_ptr, pl.d_run_base, pl.d_lpos, pl.d_rg_chunks,
                    pl.d_rg_rank, pl.d_rg_runs, pl.d_early_pos, pl.d_pe_ptr, pl.d_pe};
      for (void* p : pp)
        if (p) (void)hipFree(p);
    }
    if (!pl.owns) continue;
    void* pp[] = {pl.d_items, pl.d_ranges};
    for (void* p : pp)
      if (p) (void)hipFree(p);
  }
  if (h->h_pinned) (void)hipHostFree(h->h_pinned);
  void* rp[] = {h->d_chol, h->d_vec, h->d_dmean, h->d_info};
  for (void* p : rp)
    if (p) (void)hipFree(p);
  if (h->blas) (void)rocblas_destroy_handle(h->blas);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

bool check_handle(LeviathanHandle* h) {
  g_last_error.clear();
  if (!h) {
    set_err("null LeviathanHandle");
    return false;
  }
  if (hipSetDevice(h->device) != hipSuccess) {
    set_err("hipSetDevice(%d) failed", h->device);
    return false;
  }
  return true;
}

StepArgs step_args(LeviathanHandle* h, float dt, const DevPlan& pl) {
  StepArgs a{};
  a.row_ptr = h->d_row_ptr;
  a.col = h->d_col;
  a.delay = h->d_delay;
  a.w = h->d_w;
  a.theta = h->d_theta;
  a.theta_hat = h->d_theta_hat;
  a.omega = h->d_omega;
  a.omega_hat = h->d_omega_hat;
  a.ring = h->d_ring;
  a.ring_pitch = h->ring_pitch;
  a.n_slots = h->n_slots;
  a.num_items = pl.n_items;
  for (int x = 0; x <= lev::kNumXcd; ++x) a.xcd_split[x] = pl.xcd_split[x];
  a.items = pl.d_items;
  a.partials = h->d_partials;
  a.ctrl = h->d_ctrl;
  a.step = h->d_step;
  a.drive_b = h->drive_on ? h->d_drive_b : nullptr;
  a.drive_u = h->d_drive_u;
  a.hbuf = pl.kind == LEV_SCHEDULE_STAGED ? h->d_hbuf2 : h->d_hbuf;
  a.hstride = 1;
  a.hpos = h->d_hpos;
  a.lpos = pl.d_lpos;
  a.hbuf_next = nullptr;
  a.pair = h->d_pair;
  a.early_ptr = nullptr;
  a.early = nullptr;
  a.early_k = 0;
  a.dt = dt;
  a.eta = h->cfg.eta;
  a.lam = h->cfg.lam;
  a.gamma = h->cfg.gamma;
  a.alpha = h->cfg.alpha_eps;
  a.kappa_w = h->cfg.kappa_w;
  a.w_min = h->cfg.w_min;
  a.w_max = h->cfg.w_max;
  return a;
}

FinalizeArgs fin_args(LeviathanHandle* h, float dt, const DevPlan& pl) {
  FinalizeArgs f{};
  f.partials = h->d_partials;
  f.ctrl = h->d_ctrl;
  f.step = h->d_step;
  f.r_hist = h->d_r_hist;
  f.ranges = pl.d_ranges;
  f.n_nets = h->n_nets;
  f.num_blocks = pl.n_items;
  f.dt = dt;
  f.theta = h->d_theta;
  if (h->phi) f.phi = lev::phi_view(h->phi);
  f.chunks = h->d_chunks;
  f.fin_done = h->d_fin_done;
  return f;
}

lev::ScatterArgs scatter_args(LeviathanHandle* h, const DevPlan& pl) {
  lev::ScatterArgs sa{};
  sa.csc_ptr = h->d_csc_ptr;
  sa.delay = h->d_delay_csc;
  sa.dsrc = h->d_dsrc;
  sa.pos = pl.d_pos;
  sa.rank_run = pl.d_rank_run;
  sa.run_ptr = pl.d_run_ptr;
  sa.run_base = pl.d_run_base;
  sa.ring = h->d_ring;
  sa.ring_pitch = h->ring_pitch;
  sa.ring_bytes = (int64_t)h->ring_bytes;
  sa.n_slots = h->n_slots;
  sa.num_blocks = h->n_sources;
  sa.xcd_remap = 1;
  sa.blocks = h->d_sources;
  sa.hbuf = h->d_hbuf;
  sa.steps = 1;
  sa.step = h->d_step;
  return sa;
}

// Whether blocks of steps run paired row passes: asked for (row_steps = 2),
// possible (no delay-0 edge, so step t+1 reads nothing step t writes), and the
// block pairs up (the direct schedule, or an even number of steps per scatter
// pass).
bool paired(const LeviathanHandle* h) {
  if (h->row_steps != 2 || h->min_delay < 1 || !h->d_pair) return false;
  return h->plans[h->active].kind == LEV_SCHEDULE_DIRECT || h->scatter_steps % 2 == 0;
}

// Steps of one block under the active plan: the steps one scatter pass
// gathers (two-phase schedules), or for the direct schedule 2 with paired row
// passes and 1 without.
int block_steps(const LeviathanHandle* h) {
  if (h->plans[h->active].kind == LEV_SCHEDULE_DIRECT) return paired(h) ? 2 : 1;
  return h->scatter_steps;
}

// Enqueue the step kernel(s) of the active plan, without the finalize, for
// step k of a block of K consecutive steps (K = 1: a step on its own).
// Two-phase schedules: step 0 of a block runs the scatter (and, staged, the
// regroup) for all K steps; step k > 0 first gathers its edges with delay < k,
// straight into the array the row pass reads.
// Paired row passes (paired(h), K >= 2): an even step k runs the pair's row
// pass, which leaves the odd step's integration to pair_finish_kernel, run
// as step k + 1.  The early gathers of both steps of a pair run before it:
// without delay-0 edges the odd step's read theta(t + k) at the latest, which
// the previous pair wrote.
// mid / mid2 (optional) are recorded after the scatter phase and after the regroup.
hipError_t enqueue_main(LeviathanHandle* h, float dt, int k, int K, hipEvent_t mid = nullptr,
                        hipEvent_t mid2 = nullptr) {
  const DevPlan& pl = h->plans[h->active];
  const bool pair = K >= 2 && paired(h);
  const int gather = pl.kind == LEV_SCHEDULE_DIRECT ? 0 : pl.kind == LEV_SCHEDULE_STAGED ? 2 : 1;
  hipError_t e = hipSuccess;
  if (pair && k % 2 == 1) {  // the second step of a pair: its integration only
    if (mid) e = hipEventRecord(mid, h->stream);
    if (e == hipSuccess && mid2) e = hipEventRecord(mid2, h->stream);
    if (e == hipSuccess) e = lev::launch_pair_finish(h->cfg.history, h->cfg.layout, step_args(h, dt, pl), h->stream);
    return e;
  }
  if (pl.kind == LEV_SCHEDULE_DIRECT) {
    const StepArgs a = step_args(h, dt, pl);
    return pair ? lev::launch_pair(h->cfg.history, h->cfg.layout, pl.ept, 0, a, h->stream)
                : lev::launch_step(h->cfg.history, h->cfg.layout, pl.ept, 0, a, h->stream);
  }
  const bool staged = pl.kind == LEV_SCHEDULE_STAGED;
  const size_t hsize = h->cfg.history == LEV_HISTORY_F32 ? 4 : 2;
  // uint16 values regrouped for paired row passes leave as pairs: step 2p's and 2p + 1's values of
  // a staged slot side by side in pair array p, one 4-byte load for both steps of a pair
  const bool packed = staged && pair && h->cfg.history == LEV_HISTORY_U16;
  // the values step k's row pass reads: the regrouped array of step k (staged; packed: the first
  // of its pair's values), or component k of the bucketed slots, which hold the K steps' values
  // side by side (binned)
  auto values_of = [&](int kk) {
    const size_t at = packed ? (size_t)(kk / 2) * 2 * h->E + (size_t)(kk % 2) : (size_t)kk * h->E;
    return staged ? static_cast<void*>(static_cast<char*>(h->d_hbuf2) + at * hsize)
                  : static_cast<void*>(static_cast<char*>(h->d_hbuf) + (size_t)kk * hsize);
  };
  void* hk = values_of(k);
  const int32_t hstride = staged ? 1 : K;
  // the second pair of a staged block gathers its early edges in the row pass itself
  const bool pair_early = pair && staged && k > 0;
  if (k == 0) {
    lev::ScatterArgs sa = scatter_args(h, pl);
    sa.steps = K;
    e = lev::launch_scatter(h->cfg.history, h->cfg.layout, sa, h->stream);
  } else if (!pair_early) {
    lev::EarlyArgs ea{};
    ea.src = h->d_early_src;
    ea.pos = pl.d_early_pos;
    ea.delay = h->d_early_delay;
    ea.ring = h->d_ring;
    ea.ring_pitch = h->ring_pitch;
    ea.n_slots = h->n_slots;
    ea.hstride = hstride;
    ea.step = h->d_step;
    for (int q = 0; q < (pair ? 2 : 1) && e == hipSuccess; ++q) {
      ea.count = h->early_cnt[k + q];
      ea.hbuf = values_of(k + q);
      ea.step_off = q;
      e = lev::launch_early(h->cfg.history, h->cfg.layout, ea, h->stream);
    }
  }
  if (e == hipSuccess && mid) e = hipEventRecord(mid, h->stream);
  if (e == hipSuccess && staged && k == 0) {
    lev::RegroupArgs ra{};
    ra.chunks = pl.d_rg_chunks;
    ra.rank = pl.d_rg_rank;
    ra.runs = pl.d_rg_runs;
    ra.src = h->d_hbuf;
    ra.dst = h->d_hbuf2;
    ra.num_chunks = pl.n_rg_chunks;
    ra.stride = h->E;
    ra.steps = K;
    ra.pack_pairs = packed ? 1 : 0;
    e = lev::launch_regroup(h->cfg.history, ra, h->stream);
  }
  if (e == hipSuccess && mid2) e = hipEventRecord(mid2, h->stream);
  if (e == hipSuccess) {
    StepArgs a = step_args(h, dt, pl);
    a.hbuf = hk;
    a.hstride = hstride;
    if (pair) {
      a.hbuf_next = values_of(k + 1);
      if (pair_early) {
        a.early_ptr = pl.d_pe_ptr;
        a.early = pl.d_pe;
        a.early_k = k;
      }
      e = lev::launch_pair(h->cfg.history, h->cfg.layout, pl.ept, packed ? 3 : gather, a, h->stream);
    } else {
      e = lev::launch_step(h->cfg.history, h->cfg.layout, pl.ept, gather, a, h->stream);
    }
  }
  return e;
}

// Enqueue one step on the handle's stream (plus the feature capture while
// leviathan_step_n_features runs): step k of a block of K.
bool enqueue_step(LeviathanHandle* h, float dt, int k = 0, int K = 1) {
  LEV_HIP(enqueue_main(h, dt, k, K));
  LEV_HIP(lev::launch_finalize(fin_args(h, dt, h->plans[h->active]), h->stream));
  if (h->capturing)
    LEV_HIP(lev::launch_capture(h->d_theta, h->d_feat_nodes, h->feat_n, h->d_X, h->feat_max_rows, h->d_step,
                                h->d_feat_base, h->stream));
  return true;
}

// Enqueue n steps: whole blocks of block_steps(h) steps, then single steps.
bool enqueue_steps(LeviathanHandle* h, float dt, int n) {
  const int K = block_steps(h);
  int t = 0;
  for (; n - t >= K; t += K)
    for (int k = 0; k < K; ++k)
      if (!enqueue_step(h, dt, k, K)) return false;
  for (; t < n; ++t)
    if (!enqueue_step(h, dt)) return false;
  return true;
}

bool get_graph(LeviathanHandle* h, float dt, int chunk, hipGraphExec_t* out) {
  uint32_t bits;
  std::memcpy(&bits, &dt, 4);
  for (auto& g : h->graphs)
    if (g.dt_bits == bits && g.chunk == chunk && g.capture == h->capturing) {
      *out = g.exec;
      return true;
    }
  hipGraph_t graph = nullptr;
  LEV_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  bool ok = true;
  ok = enqueue_steps(h, dt, chunk);
  hipError_t e = hipStreamEndCapture(h->stream, &graph);
  if (!ok) {
    if (graph) (void)hipGraphDestroy(graph);
    return false;
  }
  if (e != hipSuccess) {
    set_err("hipStreamEndCapture failed: %s", hipGetErrorString(e));
    return false;
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    set_err("hipGraphInstantiate failed: %s", hipGetErrorString(e));
    return false;
  }
  if (h->graphs.size() >= 8) drop_graphs(h);
  h->graphs.push_back({bits, chunk, h->capturing, exec});
  *out = exec;
  return true;
}

// Run n (<= kRHistSteps) steps and copy their r values out.
bool run_segment(LeviathanHandle* h, float dt, int n, float* r_out) {
  const int chunk = h->cfg.graph_chunk;
  int done = 0;
  if (chunk > 0) {
    while (n - done >= chunk) {
      hipGraphExec_t exec;
      if (!get_graph(h, dt, chunk, &exec)) return false;
      LEV_HIP(hipGraphLaunch(exec, h->stream));
      done += chunk;
    }
  }
  if (done < n && !enqueue_steps(h, dt, n - done)) return false;
  const int64_t first = h->step;
  h->step += n;
  if (r_out) {
    const int nn = h->n_nets;
    int64_t slot = first % lev::kRHistSteps;
    int left = n, off = 0;
    while (left > 0) {
      const int run = (int)std::min<int64_t>(left, lev::kRHistSteps - slot);
      LEV_HIP(hipMemcpyAsync(h->h_pinned + (size_t)off * nn, h->d_r_hist + (size_t)slot * nn,
                             sizeof(float) * (size_t)run * nn, hipMemcpyDeviceToHost, h->stream));
      left -= run;
      off += run;
      slot = 0;
    }
  }
  LEV_HIP(hipStreamSynchronize(h->stream));
  if (r_out) std::memcpy(r_out, h->h_pinned, sizeof(float) * (size_t)n * h->n_nets);
  return true;
}

bool upload_drive_u(LeviathanHandle* h, const float* u, int n) {
  int64_t slot = h->step % lev::kDriveSteps;
  int left = n, off = 0;
  while (left > 0) {
    const int run = (int)std::min<int64_t>(left, lev::kDriveSteps - slot);
    LEV_HIP(hipMemcpyAsync(h->d_drive_u + slot, u + off, sizeof(float) * run, hipMemcpyHostToDevice,
                           h->stream));
    left -= run;
    off += run;
    slot = 0;
  }
  return true;
}

bool validate_inputs(int N, int max_delay, const int* row_ptr, const int* col_idx, const uint8_t* delays,
                     const float* weights, const float* theta, const float* theta_hat, const float* omega) {
  if (N <= 0) {
    set_err("N must be positive (got %d)", N);
    return false;
  }
  if (max_delay < 0 || max_delay > 255) {
    set_err("max_delay must be in [0, 255] (delays are uint8), got %d", max_delay);
    return false;
  }
  if (!row_ptr || !theta || !theta_hat || !omega) {
    set_err("null input array");
    return false;
  }
  if (row_ptr[0] != 0) {
    set_err("row_ptr[0] must be 0");
    return false;
  }
  for (int i = 0; i < N; ++i)
    if (row_ptr[i + 1] < row_ptr[i]) {
      set_err("row_ptr decreases at row %d", i);
      return false;
    }
  const int64_t E = row_ptr[N];
  if (E > 0 && (!col_idx || !delays || !weights)) {
    set_err("null edge array");
    return false;
  }
  for (int64_t e = 0; e < E; ++e) {
    if (col_idx[e] < 0 || col_idx[e] >= N) {
      set_err("col_idx[%lld] = %d out of range [0, %d)", (long long)e, col_idx[e], N);
      return false;
    }
    if (delays[e] > max_delay) {
      set_err("delays[%lld] = %d exceeds max_delay %d", (long long)e, (int)delays[e], max_delay);
      return false;
    }
  }
  return true;
}

// Share of (a sample of up to ~2^20) edges whose source lies within 256 nodes of its row: index
// locality, as in ring lattices (Watts-Strogatz: ~0.9), against ~0 for Barabasi-Albert graphs.
double near_source_share(int N, const int* row_ptr, const int* col_idx) {
  const int64_t E = row_ptr[N];
  if (E == 0) return 1.0;
  const int64_t stride = std::max<int64_t>(1, E >> 20);
  int64_t seen = 0, near = 0;
  for (int i = 0; i < N; ++i)
    for (int64_t e = (row_ptr[i] + stride - 1) / stride * stride; e < row_ptr[i + 1]; e += stride) {
      ++seen;
      near += std::abs((int64_t)col_idx[e] - i) < 256;
    }
  return seen ? (double)near / (double)seen : 1.0;
}

LeviathanHandle* create(const LeviathanConfig* cfg_in, int N, int max_delay, const int* row_ptr,
                        const int* col_idx, const uint8_t* delays, const float* weights, const float* theta,
                        const float* theta_hat, const float* omega, int n_nets, const int* net_offsets,
                        const float* g0, const int* controller) {
  g_last_error.clear();
  LeviathanConfig cfg;
  if (cfg_in)
    cfg = *cfg_in;
  else
    leviathan_config_default(&cfg);
  if (cfg.history != LEV_HISTORY_F32 && cfg.history != LEV_HISTORY_U16) {
    set_err("unknown history format %d", cfg.history);
    return nullptr;
  }
  if (cfg.layout != LEV_LAYOUT_NODE_MAJOR && cfg.layout != LEV_LAYOUT_SLOT_MAJOR && cfg.layout != LEV_LAYOUT_AUTO) {
    set_err("unknown history layout %d", cfg.layout);
    return nullptr;
  }
  if (cfg.edges_per_thread != 4 && cfg.edges_per_thread != 8 && cfg.edges_per_thread != 16) {
    set_err("edges_per_thread must be 4, 8 or 16 (got %d)", cfg.edges_per_thread);
    return nullptr;
  }
  if (!(cfg.w_min <= cfg.w_max)) {  // the step clamps weights into [w_min, w_max]
    set_err("w_min (%g) must not exceed w_max (%g)", (double)cfg.w_min, (double)cfg.w_max);
    return nullptr;
  }
  if (cfg.graph_chunk < 0 || cfg.graph_chunk > lev::kRHistSteps) {
    set_err("graph_chunk must be in [0, %d]", lev::kRHistSteps);
    return nullptr;
  }
  if (cfg.schedule < LEV_SCHEDULE_DIRECT || cfg.schedule > LEV_SCHEDULE_STAGED) {
    set_err("unknown schedule %d", cfg.schedule);
    return nullptr;
  }
  const size_t hist_bytes = cfg.history == LEV_HISTORY_F32 ? 4 : 2;
  // Default bucket: 2^18 incoming edges (1 MB of FP32 delayed phases per step).  Measured at BA
  // N = 10^7, uint16, staged/8, four steps per pass, ms per step: 2^17 edges 1.251, 2^18 1.174,
  // 3 * 2^17 1.178, 2^19 1.207 (smaller buckets: shorter scatter runs; larger: longer regroups).
  if (cfg.bucket_edges == 0) cfg.bucket_edges = 1 << 18;
  if (cfg.schedule != LEV_SCHEDULE_DIRECT && cfg.bucket_edges < 1024) {
    set_err("bucket_edges must be at least 1024");
    return nullptr;
  }
  if (!validate_inputs(N, max_delay, row_ptr, col_idx, delays, weights, theta, theta_hat, omega)) return nullptr;
  std::vector<int> offs;
  if (n_nets <= 0) {
    set_err("n_nets must be positive");
    return nullptr;
  }
  if (net_offsets) {
    offs.assign(net_offsets, net_offsets + n_nets + 1);
  } else {
    if (n_nets != 1) {
      set_err("net_offsets is required when n_nets > 1");
      return nullptr;
    }
    offs = {0, N};
  }
  if (offs[0] != 0 || offs[n_nets] != N) {
    set_err("net_offsets must start at 0 and end at N");
    return nullptr;
  }
  for (int k = 0; k < n_nets; ++k) {
    if (offs[k + 1] <= offs[k]) {
      set_err("network %d has no nodes", k);
      return nullptr;
    }
    for (int i = offs[k]; i < offs[k + 1]; ++i)
      for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
        if (col_idx[e] < offs[k] || col_idx[e] >= offs[k + 1]) {
          set_err("edge %d of row %d crosses from network %d", e, i, k);
          return nullptr;
        }
  }

  LeviathanHandle* h = new LeviathanHandle();
  h->cfg = cfg;
  h->N = N;
  h->E = row_ptr[N];
  h->max_delay = max_delay;
  h->min_delay = h->E > 0 ? *std::min_element(delays, delays + h->E) : max_delay;
  h->n_slots = max_delay + 2;
  // Automatic ring layout: node-major for a direct-schedule handle whose graph allows paired row
  // passes and has little index locality, slot-major otherwise.  With pairs, a node's slots
  // t - d and t + 1 - d share a line of the node-major ring, one gather for both steps; where
  // gathers miss the caches that halves their lines (FP32 direct/8 at Barabasi-Albert N = 10^7:
  // 2.95 ms per step against 4.30 on the slot-major ring).  Where most sources lie near their
  // rows the gathers hit the L2 on either ring and the node-major ring's partial-line history
  // writes cost more (256 Watts-Strogatz networks, direct/4-16: 1.14 against 0.90 ms); so do
  // they without pairs (5.24 against 4.87 ms at BA N = 10^7).  The two-phase schedules' scatter
  // reads slot rows.
  if (cfg.layout == LEV_LAYOUT_AUTO) {
    bool node = cfg.schedule == LEV_SCHEDULE_DIRECT && h->min_delay >= 1;
    if (node) node = near_source_share(N, row_ptr, col_idx) < 0.5;
    cfg.layout = node ? LEV_LAYOUT_NODE_MAJOR : LEV_LAYOUT_SLOT_MAJOR;
  }
  h->cfg.layout = cfg.layout;
  h->n_nets = n_nets;
  auto fail = [&]() -> LeviathanHandle* {
    std::string msg = g_last_error;
    destroy(h);
    g_last_error = msg;
    return nullptr;
  };
  auto hip_ok = [&](hipError_t e, const char* what) {
    if (e != hipSuccess) {
      set_err("%s failed: %s", what, hipGetErrorString(e));
      return false;
    }
    return true;
  };
  if (!hip_ok(hipGetDevice(&h->device), "hipGetDevice")) return fail();
  // the allocation-free readout entry point's scratch address on this device
  // (looked up once per device; a failure here is reported by that entry point)
  if (!scratch_ptr(h->device)) g_last_error.clear();
  if (!hip_ok(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate")) return fail();

  const size_t hsize = cfg.history == LEV_HISTORY_F32 ? 4 : 2;
  if (cfg.layout == LEV_LAYOUT_NODE_MAJOR) {
    const int64_t per_line = 128 / (int64_t)hsize;  // pad each node's slots to whole 128-byte lines
    h->ring_pitch = ((h->n_slots + per_line - 1) / per_line) * per_line;
    if (h->n_slots <= per_line / 2) h->ring_pitch = h->n_slots;  // tiny rings: no padding
  } else {
    h->ring_pitch = ((int64_t)N + 63) / 64 * 64;
  }
  const size_t ring_elems = cfg.layout == LEV_LAYOUT_NODE_MAJOR ? (size_t)N * h->ring_pitch
                                                               : (size_t)h->n_slots * h->ring_pitch;
  h->ring_bytes = ring_elems * hsize;  // (allocated with kScatterWinCols elements more: the scatter's window
                                       // of the last slot row reads past the ring's end)

  // work decomposition: the configured direct plan always, the other direct
  // plans and the binned plan when a schedule other than DIRECT is asked for
  h->h_ctrl.assign(n_nets, NetCtrl{});
  std::string perr;
  lev::HostPlan hp[3];
  const int epts[3] = {4, 8, 16};
  for (int q = 0; q < 3; ++q) {
    if (cfg.schedule == LEV_SCHEDULE_DIRECT && epts[q] != cfg.edges_per_thread) continue;
    if (!lev::build_direct_plan(row_ptr, n_nets, offs.data(), lev::kBlock * epts[q], hp[q], perr)) {
      set_err("%s", perr.c_str());
      return fail();
    }
    h->plans[q].built = true;
    h->plans[q].kind = LEV_SCHEDULE_DIRECT;
    h->plans[q].ept = epts[q];
    h->plans[q].n_items = (int32_t)hp[q].items.size();
    for (int x = 0; x <= lev::kNumXcd; ++x) h->plans[q].xcd_split[x] = hp[q].xcd_split[x];
  }
  // two-phase plans: the scatter plan over sources, and per plan the value slots
  lev::HostScatter hsc;
  std::vector<uint32_t> bin_pos, bin_hpos, st_pos[3];
  std::vector<uint16_t> bin_rank, st_rank[3], st_lpos[3];
  lev::HostScatterRuns bin_runs, st_runs[3];
  lev::HostRegroup st_rg[3];
  std::vector<int32_t> early_src, early_cnt;
  std::vector<uint8_t> early_delay;
  std::vector<uint32_t> early_csr, bin_early_pos, st_early_pos[3];
  std::vector<int32_t> st_pe_ptr[3];
  std::vector<int2> st_pe[3];
  const bool want_binned = cfg.schedule == LEV_SCHEDULE_BINNED || cfg.schedule == LEV_SCHEDULE_AUTO;
  // the scatter kernel's launcher makes the same choice (leviathan_binned.hip)
  const bool scatter_window = lev::scatter_uses_window(cfg.layout, h->n_slots, hsize);
  const bool want_staged = cfg.schedule == LEV_SCHEDULE_STAGED || cfg.schedule == LEV_SCHEDULE_AUTO;
  if (want_binned || want_staged) {
    if (!lev::build_scatter_plan(N, row_ptr, col_idx, delays, hsc, perr)) {
      set_err("%s", perr.c_str());
      return fail();
    }
    h->n_sources = (int32_t)hsc.sources.size();
    if (cfg.layout == LEV_LAYOUT_SLOT_MAJOR) lev::order_scatter_by_delay(hsc);
    lev::early_edges(N, row_ptr, col_idx, delays, lev::kMaxScatterSteps - 1, early_src, early_delay, early_csr,
                     early_cnt);
    for (int k = 0; k < lev::kMaxScatterSteps; ++k) h->early_cnt[k] = early_cnt[k];
  }
  if (want_binned) {
    if (!lev::bucket_positions(N, row_ptr, col_idx, cfg.bucket_edges, bin_hpos, h->n_buckets, perr) ||
        !lev::csc_positions(hsc, bin_hpos, bin_pos, perr)) {
      set_err("%s", perr.c_str());
      return fail();
    }
    lev::sort_scatter_writes(hsc, bin_pos, bin_rank);
    lev::encode_scatter_runs(hsc, bin_pos, bin_runs);
    bin_early_pos.resize(early_csr.size());
    for (size_t i = 0; i < early_csr.size(); ++i) bin_early_pos[i] = bin_hpos[early_csr[i]];
    for (int q = 0; q < 3; ++q) {
      DevPlan& pb = h->plans[kPlanBinned + q];
      pb.built = true;
      pb.kind = LEV_SCHEDULE_BINNED;
      pb.ept = epts[q];
      pb.n_items = h->plans[q].n_items;
      for (int x = 0; x <= lev::kNumXcd; ++x) pb.xcd_split[x] = h->plans[q].xcd_split[x];
      pb.owns = false;
    }
  }
  if (want_staged) {
    for (int q = 0; q < 3; ++q) {
      // the scatter fills buckets of whole row blocks; the regroup moves each value to its staged slot
      std::vector<uint32_t> hpos, bhpos;
      std::vector<int32_t> bucket_begin;
      if (!lev::staged_positions(N, row_ptr, col_idx, hp[q], hpos, st_lpos[q], perr) ||
          !lev::staged_bucket_positions(N, row_ptr, col_idx, hp[q], cfg.bucket_edges, bhpos, bucket_begin, perr) ||
          !lev::csc_positions(hsc, bhpos, st_pos[q], perr) ||
          !lev::build_regroup(bhpos, hpos, bucket_begin, st_rg[q], perr)) {
        set_err("%s", perr.c_str());
        return fail();
      }
      lev::sort_scatter_writes(hsc, st_pos[q], st_rank[q]);
      lev::encode_scatter_runs(hsc, st_pos[q], st_runs[q]);
      // early edges are gathered straight into the regrouped array: their staged slots
      st_early_pos[q].resize(early_csr.size());
      for (size_t i = 0; i < early_csr.size(); ++i) st_early_pos[q][i] = hpos[early_csr[i]];
      if (!lev::block_early_edges(hp[q], early_src, early_delay, early_csr, hpos, st_pe_ptr[q], st_pe[q], perr)) {
        set_err("%s", perr.c_str());
        return fail();
      }
      DevPlan& ps = h->plans[kPlanStaged + q];
      ps.built = true;
      ps.kind = LEV_SCHEDULE_STAGED;
      ps.ept = epts[q];
      ps.n_items = h->plans[q].n_items;
      for (int x = 0; x <= lev::kNumXcd; ++x) ps.xcd_split[x] = h->plans[q].xcd_split[x];
      ps.owns = false;
    }
  }
  hsc.csc_edge = std::vector<uint32_t>();
  h->active = plan_base(cfg.schedule == LEV_SCHEDULE_AUTO ? LEV_SCHEDULE_DIRECT : cfg.schedule) +
              plan_of_ept(cfg.edges_per_thread);
  for (const auto& pl : h->plans)
    if (pl.built) h->n_partials = std::max(h->n_partials, pl.n_items);
  for (int k = 0; k < n_nets; ++k) {
    NetCtrl& c = h->h_ctrl[k];
    c.g = g0 ? (double)g0[k] : (double)cfg.g0;
    c.r = 0.0;
    c.r0 = cfg.r0;
    c.beta = cfg.beta;
    c.mu = cfg.mu;
    c.g_rest = cfg.g_rest;
    c.g_min = cfg.g_min;
    c.g_max = cfg.g_max;
    c.controller = controller ? (controller[k] != 0) : (cfg.gain_controller != 0);
    c.n_nodes = offs[k + 1] - offs[k];
  }

  if (!dalloc(h, &h->d_row_ptr, (size_t)N + 1) || !dalloc(h, &h->d_col, (size_t)h->E) ||
      !dalloc(h, &h->d_delay, (size_t)h->E) || !dalloc(h, &h->d_w, (size_t)h->E) ||
      !dalloc(h, &h->d_theta, (size_t)N) || !dalloc(h, &h->d_theta_hat, (size_t)N) ||
      !dalloc(h, &h->d_omega, (size_t)N) || !dalloc(h, &h->d_omega_hat, (size_t)N) ||
      !dalloc(h, (char**)&h->d_ring, (ring_elems + lev::kScatterWinCols) * hsize) ||
      !dalloc(h, &h->d_partials, (size_t)h->n_partials) ||
      !dalloc(h, &h->d_chunks, (size_t)(h->n_partials + lev::kChunkPartials - 1) / lev::kChunkPartials) ||
      !dalloc(h, &h->d_fin_done, 0) ||
      !dalloc(h, &h->d_ctrl, (size_t)n_nets) ||
      !dalloc(h, &h->d_step, 1) || !dalloc(h, &h->d_r_hist, (size_t)lev::kRHistSteps * n_nets) ||
      !dalloc(h, &h->d_scratch, (size_t)N) || !dalloc(h, &h->d_drive_u, (size_t)lev::kDriveSteps) ||
      (h->min_delay >= 1 && !dalloc(h, &h->d_pair, (size_t)N)))
    return fail();
  // paired row passes by default wherever the graph allows them (bit-identical to single steps)
  h->row_steps = h->min_delay >= 1 ? 2 : 1;
  if (!hip_ok(hipHostMalloc((void**)&h->h_pinned, sizeof(float) * (size_t)lev::kRHistSteps * n_nets,
                            hipHostMallocDefault),
              "hipHostMalloc"))
    return fail();

  const hipMemcpyKind H2D = hipMemcpyHostToDevice;
  const int64_t zero = 0;
  bool ok = hip_ok(hipMemcpy(h->d_row_ptr, row_ptr, sizeof(int) * ((size_t)N + 1), H2D), "copy row_ptr") &&
            (h->E == 0 || (hip_ok(hipMemcpy(h->d_col, col_idx, sizeof(int) * h->E, H2D), "copy col_idx") &&
                           hip_ok(hipMemcpy(h->d_delay, delays, h->E, H2D), "copy delays") &&
                           hip_ok(hipMemcpy(h->d_w, weights, sizeof(float) * h->E, H2D), "copy weights"))) &&
            hip_ok(hipMemcpy(h->d_theta, theta, sizeof(float) * N, H2D), "copy theta") &&
            hip_ok(hipMemcpy(h->d_theta_hat, theta_hat, sizeof(float) * N, H2D), "copy theta_hat") &&
            hip_ok(hipMemcpy(h->d_omega, omega, sizeof(float) * N, H2D), "copy omega") &&
            // omega_hat starts as omega (leviathan_kernel.cu:350)
            hip_ok(hipMemcpy(h->d_omega_hat, omega, sizeof(float) * N, H2D), "copy omega_hat") &&
            hip_ok(hipMemcpy(h->d_ctrl, h->h_ctrl.data(), sizeof(NetCtrl) * n_nets, H2D), "copy ctrl") &&
            hip_ok(hipMemcpy(h->d_step, &zero, sizeof(int64_t), H2D), "copy step") &&
            hip_ok(hipMemset(h->d_r_hist, 0, sizeof(float) * (size_t)lev::kRHistSteps * n_nets), "memset r") &&
            hip_ok(hipMemset(h->d_drive_u, 0, sizeof(float) * lev::kDriveSteps), "memset u") &&
            hip_ok(hipMemset(h->d_partials, 0, sizeof(double2) * std::max(h->n_partials, 1)), "memset partials");
  if (!ok) return fail();
  for (int q = 0; q < 3; ++q) {
    if (!h->plans[q].built) continue;
    DevPlan& pl = h->plans[q];
    if (!dalloc(h, &pl.d_items, hp[q].items.size()) || !dalloc(h, &pl.d_ranges, (size_t)n_nets)) return fail();
    if (!hip_ok(hipMemcpy(pl.d_items, hp[q].items.data(), sizeof(int4) * hp[q].items.size(), H2D), "copy plan") ||
        !hip_ok(hipMemcpy(pl.d_ranges, hp[q].ranges.data(), sizeof(int4) * n_nets, H2D), "copy ranges"))
      return fail();
  }
  if (want_binned || want_staged) {
    const size_t E = (size_t)h->E;
    if (!dalloc(h, &h->d_sources, hsc.sources.size()) || !dalloc(h, &h->d_csc_ptr, (size_t)N + 1) ||
        !dalloc(h, &h->d_delay_csc, E) || !dalloc(h, (char**)&h->d_hbuf, E * hist_bytes) ||
        (want_staged && !dalloc(h, (char**)&h->d_hbuf2, E * hist_bytes)))
      return fail();
    ok = hip_ok(hipMemcpy(h->d_sources, hsc.sources.data(), sizeof(int4) * hsc.sources.size(), H2D),
                "copy sources") &&
         hip_ok(hipMemcpy(h->d_csc_ptr, hsc.csc_ptr.data(), sizeof(int32_t) * ((size_t)N + 1), H2D), "copy csc") &&
         (E == 0 || hip_ok(hipMemcpy(h->d_delay_csc, hsc.delay_csc.data(), E, H2D), "copy csc delays"));
    if (ok && cfg.layout == LEV_LAYOUT_SLOT_MAJOR && !scatter_window)
      ok = dalloc(h, &h->d_dsrc, E) &&
           (E == 0 || hip_ok(hipMemcpy(h->d_dsrc, hsc.dsrc.data(), sizeof(uint16_t) * E, H2D), "copy edge sources"));
    const size_t ne = early_src.size();
    ok = ok && dalloc(h, &h->d_early_src, ne) && dalloc(h, &h->d_early_delay, ne) &&
         (ne == 0 || (hip_ok(hipMemcpy(h->d_early_src, early_src.data(), sizeof(int32_t) * ne, H2D), "copy early") &&
                      hip_ok(hipMemcpy(h->d_early_delay, early_delay.data(), ne, H2D), "copy early")));
    if (!ok) return fail();
    if (want_binned) {
      if (!dalloc(h, &h->d_pos, E) || !dalloc(h, &h->d_hpos, E) ||
          !upload_runs(h, bin_rank, bin_runs, &h->d_rank_run, &h->d_run_ptr, &h->d_run_base,
                       scatter_window ? &hsc : nullptr) ||
          !dalloc(h, &h->d_early_pos, ne))
        return fail();
      if (ne && !hip_ok(hipMemcpy(h->d_early_pos, bin_early_pos.data(), sizeof(uint32_t) * ne, H2D), "copy early"))
        return fail();
      if (E && !(hip_ok(hipMemcpy(h->d_pos, bin_pos.data(), sizeof(uint32_t) * E, H2D), "copy positions") &&
                 hip_ok(hipMemcpy(h->d_hpos, bin_hpos.data(), sizeof(uint32_t) * E, H2D), "copy slots")))
        return fail();
      for (int q = 0; q < 3; ++q) {
        DevPlan& pb = h->plans[kPlanBinned + q];
        pb.d_items = h->plans[q].d_items;
        pb.d_ranges = h->plans[q].d_ranges;
        pb.d_pos = h->d_pos;
        pb.d_rank_run = h->d_rank_run;
        pb.d_run_ptr = h->d_run_ptr;
        pb.d_run_base = h->d_run_base;
        pb.d_early_pos = h->d_early_pos;
      }
    }
    if (want_staged) {
      for (int q = 0; q < 3; ++q) {
        DevPlan& ps = h->plans[kPlanStaged + q];
        ps.d_items = h->plans[q].d_items;
        ps.d_ranges = h->plans[q].d_ranges;
        ps.owns_pos = true;
        const lev::HostRegroup& rg = st_rg[q];
        ps.n_rg_chunks = (int32_t)rg.chunks.size();
        if (!dalloc(h, &ps.d_pos, E) || !dalloc(h, &ps.d_lpos, E) ||
            !upload_runs(h, st_rank[q], st_runs[q], &ps.d_rank_run, &ps.d_run_ptr, &ps.d_run_base,
                         scatter_window ? &hsc : nullptr) ||
            !dalloc(h, &ps.d_rg_chunks, rg.chunks.size()) || !dalloc(h, &ps.d_rg_rank, E) ||
            !dalloc(h, &ps.d_rg_runs, rg.runs.size()) || !dalloc(h, &ps.d_early_pos, ne) ||
            !dalloc(h, &ps.d_pe_ptr, st_pe_ptr[q].size()) || !dalloc(h, &ps.d_pe, st_pe[q].size()))
          return fail();
        if (!hip_ok(hipMemcpy(ps.d_pe_ptr, st_pe_ptr[q].data(), sizeof(int32_t) * st_pe_ptr[q].size(), H2D),
                    "copy block early edges") ||
            (!st_pe[q].empty() &&
             !hip_ok(hipMemcpy(ps.d_pe, st_pe[q].data(), sizeof(int2) * st_pe[q].size(), H2D), "copy block early edges")))
          return fail();
        if (ne && !hip_ok(hipMemcpy(ps.d_early_pos, st_early_pos[q].data(), sizeof(uint32_t) * ne, H2D),
                          "copy early"))
          return fail();
        if (E && !(hip_ok(hipMemcpy(ps.d_pos, st_pos[q].data(), sizeof(uint32_t) * E, H2D), "copy positions") &&
                   hip_ok(hipMemcpy(ps.d_lpos, st_lpos[q].data(), sizeof(uint16_t) * E, H2D), "copy slots") &&
                   hip_ok(hipMemcpy(ps.d_rg_chunks, rg.chunks.data(), sizeof(int4) * rg.chunks.size(), H2D),
                          "copy regroup chunks") &&
                   hip_ok(hipMemcpy(ps.d_rg_rank, rg.rank.data(), sizeof(uint16_t) * E, H2D), "copy regroup ranks") &&
                   hip_ok(hipMemcpy(ps.d_rg_runs, rg.runs.data(), sizeof(int2) * rg.runs.size(), H2D),
                          "copy regroup runs")))
          return fail();
      }
    }
  }
  // history pre-fill (init_buffer_kernel, leviathan_kernel.cu:371-372)
  if (!hip_ok(lev::launch_fill_ring(cfg.history, cfg.layout, h->d_ring, h->ring_pitch, h->n_slots, h->d_theta, N,
                                    h->stream),
              "fill history") ||
      !hip_ok(hipStreamSynchronize(h->stream), "hipStreamSynchronize"))
    return fail();
  return h;
}

bool refresh_ctrl_field(LeviathanHandle* h, int net, size_t offset, const void* src, size_t bytes) {
  LEV_HIP(hipStreamSynchronize(h->stream));
  LEV_HIP(hipMemcpy(reinterpret_cast<char*>(h->d_ctrl + net) + offset, src, bytes, hipMemcpyHostToDevice));
  return true;
}

}  // namespace

namespace {

// The FP32 matrix-core Gram of the centred feature rows [row0, row0 + nrows)
// into a new device array *d_g (F x F, the caller frees it).
bool gram_on_device(LeviathanHandle* h, int64_t row0, int64_t nrows, const double* mean, int terms, float** d_g,
                    double* ms_out) {
  const int32_t F = 2 * h->feat_n;
  const int32_t Fp = (F + 255) / 256 * 256;  // whole 256-column Gram tiles
  const int64_t Tp = (nrows + 63) / 64 * 64;
  double* d_mean = nullptr;
  uint16_t *d_xct = nullptr, *d_lo = nullptr;
  float* g = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc((void**)&d_mean, sizeof(double) * F);
  if (e == hipSuccess) e = hipMalloc((void**)&d_xct, sizeof(uint16_t) * (size_t)Fp * Tp);
  if (e == hipSuccess && terms == 3) e = hipMalloc((void**)&d_lo, sizeof(uint16_t) * (size_t)Fp * Tp);
  if (e == hipSuccess) e = hipMalloc((void**)&g, sizeof(float) * (size_t)F * F);
  if (e == hipSuccess) e = hipMemcpy(d_mean, mean, sizeof(double) * F, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = lev::launch_center_bf16(h->d_X, F, row0, nrows, d_mean, F, Fp, d_xct, Tp, d_lo, h->stream);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
  if (e == hipSuccess)
    e = terms == 1 ? lev::launch_gram(d_xct, Tp, Fp, F, g, h->stream)
                   : lev::launch_gram_split(d_xct, d_lo, Tp, Fp, F, g, h->stream);
  if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && ms_out) {
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, e0, e1);
    *ms_out = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  void* ptrs[] = {d_mean, d_xct, d_lo};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (e != hipSuccess) {
    if (g) (void)hipFree(g);
    set_err("features Gram: %s", hipGetErrorString(e));
    return false;
  }
  *d_g = g;
  return true;
}

}  // namespace

namespace lev {
void error_set(const char* msg) { g_last_error = msg; }
void error_clear() { g_last_error.clear(); }
bool error_empty() { return g_last_error.empty(); }
}  // namespace lev

extern "C" {

const char* leviathan_last_error(void) { return g_last_error.c_str(); }

void leviathan_config_default(LeviathanConfig* c) {
  if (!c) return;
  // leviathan_kernel.cu:281-297; g_rest is the literal 1.5f of :433
  c->eta = 0.01f;
  c->lam = 0.001f;
  c->gamma = 0.05f;
  c->alpha_eps = 0.1f;
  c->kappa_w = 0.01f;
  c->w_min = 0.0f;
  c->w_max = 2.0f;
  c->g0 = 1.5f;
  c->r0 = 0.5f;
  c->beta = 0.01f;
  c->mu = 0.01f;
  c->g_rest = 1.5f;
  c->g_min = 0.1f;
  c->g_max = 5.0f;
  c->gain_controller = 1;
  c->history = LEV_HISTORY_F32;
  c->layout = LEV_LAYOUT_AUTO;
  c->edges_per_thread = 8;
  c->graph_chunk = 100;
  c->schedule = LEV_SCHEDULE_DIRECT;
  c->bucket_edges = 0;
}

LeviathanHandle* leviathan_init(int N, int max_delay, const int* h_row_ptr, const int* h_col_idx,
                                const uint8_t* h_delays, const float* h_weights, const float* h_theta,
                                const float* h_theta_hat, const float* h_omega) {
  return create(nullptr, N, max_delay, h_row_ptr, h_col_idx, h_delays, h_weights, h_theta, h_theta_hat, h_omega, 1,
                nullptr, nullptr, nullptr);
}

LeviathanHandle* leviathan_init_ex(const LeviathanConfig* cfg, int N, int max_delay, const int* h_row_ptr,
                                   const int* h_col_idx, const uint8_t* h_delays, const float* h_weights,
                                   const float* h_theta, const float* h_theta_hat, const float* h_omega, int n_nets,
                                   const int* net_offsets, const float* g0, const int* controller) {
  return create(cfg, N, max_delay, h_row_ptr, h_col_idx, h_delays, h_weights, h_theta, h_theta_hat, h_omega, n_nets,
                net_offsets, g0, controller);
}

float leviathan_step(LeviathanHandle* h, float dt) {
  if (!check_handle(h)) return std::numeric_limits<float>::quiet_NaN();
  if (h->drive_on) {
    // a drive without an input sequence holds u = 0
    const float zero = 0.0f;
    if (!upload_drive_u(h, &zero, 1)) return std::numeric_limits<float>::quiet_NaN();
  }
  const int64_t slot = h->step % lev::kRHistSteps;
  if (!enqueue_step(h, dt)) return std::numeric_limits<float>::quiet_NaN();
  hipError_t e = hipMemcpyAsync(h->h_pinned, h->d_r_hist + slot * h->n_nets, sizeof(float), hipMemcpyDeviceToHost,
                                h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    set_err("leviathan_step: %s", hipGetErrorString(e));
    return std::numeric_limits<float>::quiet_NaN();
  }
  h->step += 1;
  return h->h_pinned[0];
}

int leviathan_step_n(LeviathanHandle* h, float dt, int n, float* r_out) {
  if (!check_handle(h)) return -1;
  if (n < 0) {
    set_err("n must be >= 0");
    return -1;
  }
  const int seg_max = std::min(lev::kRHistSteps, lev::kDriveSteps);
  std::vector<float> zeros;
  for (int done = 0; done < n;) {
    const int seg = std::min(n - done, seg_max);
    if (h->drive_on) {
      zeros.assign(seg, 0.0f);
      if (!upload_drive_u(h, zeros.data(), seg)) return -1;
    }
    if (!run_segment(h, dt, seg, r_out ? r_out + (size_t)done * h->n_nets : nullptr)) return -1;
    done += seg;
  }
  return 0;
}

int leviathan_step_n_driven(LeviathanHandle* h, float dt, int n, const float* u, float* r_out) {
  if (!check_handle(h)) return -1;
  if (!h->drive_on) {
    set_err("no drive set: call leviathan_set_drive first");
    return -1;
  }
  if (n < 0 || (n > 0 && !u)) {
    set_err("need n >= 0 and an input sequence");
    return -1;
  }
  const int seg_max = std::min(lev::kRHistSteps, lev::kDriveSteps);
  for (int done = 0; done < n;) {
    const int seg = std::min(n - done, seg_max);
    if (!upload_drive_u(h, u + done, seg)) return -1;
    if (!run_segment(h, dt, seg, r_out ? r_out + (size_t)done * h->n_nets : nullptr)) return -1;
    done += seg;
  }
  return 0;
}

int leviathan_set_drive(LeviathanHandle* h, const float* b) {
  if (!check_handle(h)) return -1;
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    set_err("hipStreamSynchronize failed");
    return -1;
  }
  drop_graphs(h);  // captured graphs hold the drive pointer
  if (!b) {
    h->drive_on = false;
    return 0;
  }
  if (!h->d_drive_b && !dalloc(h, &h->d_drive_b, (size_t)h->N)) return -1;
  hipError_t e = hipMemcpy(h->d_drive_b, b, sizeof(float) * h->N, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_err("copy drive: %s", hipGetErrorString(e));
    return -1;
  }
  h->drive_on = true;
  return 0;
}

void leviathan_get_theta(LeviathanHandle* h, float* out) {
  if (!check_handle(h)) return;
  if (!out) {
    set_err("null output array");
    return;
  }
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(out, h->d_theta, sizeof(float) * h->N, hipMemcpyDeviceToHost);
  if (e != hipSuccess) set_err("leviathan_get_theta: %s", hipGetErrorString(e));
}

void leviathan_set_theta(LeviathanHandle* h, const float* in) {
  if (!check_handle(h)) return;
  if (!in) {
    set_err("null input array");
    return;
  }
  hipError_t e = hipMemcpyAsync(h->d_theta, in, sizeof(float) * h->N, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = lev::launch_fill_ring(h->cfg.history, h->cfg.layout, h->d_ring, h->ring_pitch, h->n_slots, h->d_theta, h->N,
                              h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) set_err("leviathan_set_theta: %s", hipGetErrorString(e));
}

void leviathan_free(LeviathanHandle* h) {
  g_last_error.clear();
  destroy(h);
}

int leviathan_set_gain_net(LeviathanHandle* h, int net, float g) {
  if (!check_handle(h)) return -1;
  if (net < 0 || net >= h->n_nets) {
    set_err("network %d out of range", net);
    return -1;
  }
  const double gd = g;
  return refresh_ctrl_field(h, net, offsetof(NetCtrl, g), &gd, sizeof(double)) ? 0 : -1;
}

float leviathan_get_gain_net(LeviathanHandle* h, int net) {
  if (!check_handle(h)) return std::numeric_limits<float>::quiet_NaN();
  if (net < 0 || net >= h->n_nets) {
    set_err("network %d out of range", net);
    return std::numeric_limits<float>::quiet_NaN();
  }
  double g = 0.0;
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess)
    e = hipMemcpy(&g, reinterpret_cast<char*>(h->d_ctrl + net) + offsetof(NetCtrl, g), sizeof(double),
                  hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_err("get_gain: %s", hipGetErrorString(e));
    return std::numeric_limits<float>::quiet_NaN();
  }
  return (float)g;
}

int leviathan_set_gain_controller_net(LeviathanHandle* h, int net, bool enabled) {
  if (!check_handle(h)) return -1;
  if (net < 0 || net >= h->n_nets) {
    set_err("network %d out of range", net);
    return -1;
  }
  const int32_t v = enabled ? 1 : 0;
  return refresh_ctrl_field(h, net, offsetof(NetCtrl, controller), &v, sizeof(int32_t)) ? 0 : -1;
}

void leviathan_set_gain(LeviathanHandle* h, float g) { leviathan_set_gain_net(h, 0, g); }

float leviathan_get_gain(LeviathanHandle* h) { return leviathan_get_gain_net(h, 0); }

void leviathan_set_gain_controller(LeviathanHandle* h, bool enabled) {
  leviathan_set_gain_controller_net(h, 0, enabled);
}

size_t leviathan_get_vram_usage(LeviathanHandle* h) {
  if (!check_handle(h)) return 0;
  return h->vram;
}

void leviathan_reset_weights(LeviathanHandle* h, const float* w) {
  if (!check_handle(h)) return;
  if (!w && h->E > 0) {
    set_err("null weights array");
    return;
  }
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && h->E > 0) e = hipMemcpy(h->d_w, w, sizeof(float) * h->E, hipMemcpyHostToDevice);
  if (e != hipSuccess) set_err("leviathan_reset_weights: %s", hipGetErrorString(e));
}

int leviathan_get_weights(LeviathanHandle* h, float* out) {
  if (!check_handle(h)) return -1;
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && h->E > 0) e = hipMemcpy(out, h->d_w, sizeof(float) * h->E, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_err("get_weights: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int leviathan_get_predictor(LeviathanHandle* h, float* theta_hat, float* omega_hat) {
  if (!check_handle(h)) return -1;
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && theta_hat)
    e = hipMemcpy(theta_hat, h->d_theta_hat, sizeof(float) * h->N, hipMemcpyDeviceToHost);
  if (e == hipSuccess && omega_hat)
    e = hipMemcpy(omega_hat, h->d_omega_hat, sizeof(float) * h->N, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_err("get_predictor: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int leviathan_get_history(LeviathanHandle* h, int delay, float* out) {
  if (!check_handle(h)) return -1;
  if (delay < 0 || delay > h->max_delay) {
    set_err("delay %d out of [0, %d]", delay, h->max_delay);
    return -1;
  }
  const int32_t slot = (int32_t)(((h->step - delay) % h->n_slots + h->n_slots) % h->n_slots);
  hipError_t e = lev::launch_read_slot(h->cfg.history, h->cfg.layout, h->d_ring, h->ring_pitch, slot, h->d_scratch,
                                       h->N, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(out, h->d_scratch, sizeof(float) * h->N, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_err("get_history: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int leviathan_set_predictor(LeviathanHandle* h, const float* theta_hat, const float* omega_hat) {
  if (!check_handle(h)) return -1;
  if (!theta_hat || !omega_hat) {
    set_err("set_predictor needs theta_hat and omega_hat");
    return -1;
  }
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(h->d_theta_hat, theta_hat, sizeof(float) * h->N, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->d_omega_hat, omega_hat, sizeof(float) * h->N, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_err("set_predictor: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int leviathan_set_history(LeviathanHandle* h, int delay, const float* values) {
  if (!check_handle(h)) return -1;
  if (delay < 1 || delay > h->max_delay || !values) {
    set_err("set_history needs values and a delay in [1, %d] (delay 0 is theta: set_theta)", h->max_delay);
    return -1;
  }
  const int32_t slot = (int32_t)(((h->step - delay) % h->n_slots + h->n_slots) % h->n_slots);
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(h->d_scratch, values, sizeof(float) * h->N, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = lev::launch_write_slot(h->cfg.history, h->cfg.layout, h->d_ring, h->ring_pitch, slot, h->d_scratch, h->N,
                               h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    set_err("set_history: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int leviathan_get_gains(LeviathanHandle* h, double* g) {
  if (!check_handle(h)) return -1;
  if (!g) {
    set_err("get_gains needs an output array");
    return -1;
  }
  std::vector<NetCtrl> ctrl((size_t)h->n_nets);
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(ctrl.data(), h->d_ctrl, sizeof(NetCtrl) * h->n_nets, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_err("get_gains: %s", hipGetErrorString(e));
    return -1;
  }
  for (int k = 0; k < h->n_nets; ++k) g[k] = ctrl[k].g;
  return 0;
}

int leviathan_set_gains(LeviathanHandle* h, const double* g) {
  if (!check_handle(h)) return -1;
  if (!g) {
    set_err("set_gains needs the gains");
    return -1;
  }
  for (int k = 0; k < h->n_nets; ++k)
    if (!refresh_ctrl_field(h, k, offsetof(NetCtrl, g), g + k, sizeof(double))) return -1;
  return 0;
}

int leviathan_set_step_count(LeviathanHandle* h, int64_t step) {
  if (!check_handle(h)) return -1;
  if (step < 0) {
    set_err("step count must be non-negative");
    return -1;
  }
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(h->d_step, &step, sizeof(int64_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_err("set_step_count: %s", hipGetErrorString(e));
    return -1;
  }
  h->step = step;
  return 0;
}

int leviathan_attach_phi(LeviathanHandle* h, PhiHandle* phi) {
  if (!check_handle(h)) return -1;
  if (phi == h->phi) return 0;
  if (phi) {
    if (lev::phi_device(phi) != h->device) {
      set_err("attach_phi: the PhiHandle lives on device %d, the engine on %d", lev::phi_device(phi), h->device);
      return -1;
    }
    if (lev::phi_max_hub(phi) >= h->N) {
      set_err("attach_phi: hub index %d out of range for %d nodes", lev::phi_max_hub(phi), h->N);
      return -1;
    }
    if (lev::phi_engine(phi)) {
      set_err("attach_phi: the PhiHandle is attached to another engine");
      return -1;
    }
  }
  const hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    set_err("attach_phi: %s", hipGetErrorString(e));
    return -1;
  }
  if (h->phi) lev::phi_link(h->phi, nullptr, nullptr);
  h->phi = phi;
  if (phi) lev::phi_link(phi, h, h->stream);
  drop_graphs(h);  // captured finalize arguments name the old handle
  return 0;
}

int leviathan_num_nodes(LeviathanHandle* h) { return h ? h->N : 0; }
int64_t leviathan_num_edges(LeviathanHandle* h) { return h ? h->E : 0; }
int64_t leviathan_step_count(LeviathanHandle* h) { return h ? h->step : 0; }
int leviathan_num_nets(LeviathanHandle* h) { return h ? h->n_nets : 0; }
int leviathan_num_blocks(LeviathanHandle* h) { return h ? h->plans[h->active].n_items : 0; }

int leviathan_profile_steps(LeviathanHandle* h, float dt, int n, double* out) {
  if (!check_handle(h)) return -1;
  if (n <= 0 || !out) {
    set_err("need n > 0 and an output array");
    return -1;
  }
  if (h->drive_on) {
    set_err("profile_steps runs undriven steps; switch the drive off first");
    return -1;
  }
  const bool binned = h->plans[h->active].kind != LEV_SCHEDULE_DIRECT;
  const bool staged = h->plans[h->active].kind == LEV_SCHEDULE_STAGED;
  // per step: [start, after scatter, after regroup (staged only), after the row pass, after finalize]
  constexpr int K = 5;
  std::vector<hipEvent_t> ev((size_t)K * n, nullptr);
  bool ok = true;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) ok = false;
  const FinalizeArgs f = fin_args(h, dt, h->plans[h->active]);
  // whole blocks of the multi-step scatter, then single steps; the scatter
  // phase of a step is the block's scatter pass (its first step) or the
  // gather of its early edges
  const int B = block_steps(h), full = n / B * B;
  for (int t = 0; t < n && ok; ++t) {
    hipEvent_t* e = &ev[(size_t)K * t];
    ok = hipEventRecord(e[0], h->stream) == hipSuccess &&
         enqueue_main(h, dt, t < full ? t % B : 0, t < full ? B : 1, binned ? e[1] : nullptr,
                      staged ? e[2] : nullptr) == hipSuccess &&
         hipEventRecord(e[3], h->stream) == hipSuccess && lev::launch_finalize(f, h->stream) == hipSuccess &&
         hipEventRecord(e[4], h->stream) == hipSuccess;
  }
  ok = ok && hipStreamSynchronize(h->stream) == hipSuccess;
  if (ok) {
    double kmain = 0.0, kfin = 0.0, kscat = 0.0, kreg = 0.0;
    for (int t = 0; t < n; ++t) {
      hipEvent_t* e = &ev[(size_t)K * t];
      float a = 0.f, b = 0.f, c = 0.f, d = 0.f;
      if (staged) {
        (void)hipEventElapsedTime(&c, e[0], e[1]);
        (void)hipEventElapsedTime(&d, e[1], e[2]);
        (void)hipEventElapsedTime(&a, e[2], e[3]);
      } else if (binned) {
        (void)hipEventElapsedTime(&c, e[0], e[1]);
        (void)hipEventElapsedTime(&a, e[1], e[3]);
      } else {
        (void)hipEventElapsedTime(&a, e[0], e[3]);
      }
      (void)hipEventElapsedTime(&b, e[3], e[4]);
      kmain += a;
      kfin += b;
      kscat += c;
      kreg += d;
    }
    float wall = 0.f;
    (void)hipEventElapsedTime(&wall, ev[0], ev[(size_t)K * n - 1]);
    out[0] = kmain / n;
    out[1] = kfin / n;
    out[2] = wall;
    out[3] = kscat / n;
    out[4] = kreg / n;
    h->step += n;
  } else {
    set_err("profile_steps: a HIP call failed");
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  return ok ? 0 : -1;
}

int leviathan_set_schedule(LeviathanHandle* h, int schedule, int ept) {
  if (!check_handle(h)) return -1;
  if ((schedule != LEV_SCHEDULE_DIRECT && schedule != LEV_SCHEDULE_BINNED && schedule != LEV_SCHEDULE_STAGED) ||
      (ept != 4 && ept != 8 && ept != 16)) {
    set_err("schedule must be DIRECT, BINNED or STAGED with edges_per_thread 4, 8 or 16");
    return -1;
  }
  const int idx = plan_base(schedule) + plan_of_ept(ept);
  if (!h->plans[idx].built) {
    set_err("that schedule was not built for this handle (create it with LEV_SCHEDULE_AUTO)");
    return -1;
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    set_err("hipStreamSynchronize failed");
    return -1;
  }
  if (idx != h->active) drop_graphs(h);  // captured graphs hold the old plan's arguments
  h->active = idx;
  return 0;
}

int leviathan_get_schedule(LeviathanHandle* h, int* schedule, int* ept) {
  if (!check_handle(h)) return -1;
  if (schedule) *schedule = h->plans[h->active].kind;
  if (ept) *ept = h->plans[h->active].ept;
  return 0;
}

namespace {

// Copy of everything a step changes, for trial steps that are rolled back.
struct StateSnapshot {
  struct Buf {
    void* dev;
    size_t bytes;
    void* save;
  };
  std::vector<Buf> bufs;
  int64_t step = 0;
  PhiHandle* phi = nullptr;
  LeviathanHandle* h = nullptr;

  // Takes the snapshot and detaches the hub statistics (trial steps record nothing).
  bool take(LeviathanHandle* hh) {
    h = hh;
    bufs = {{h->d_w, sizeof(float) * (size_t)h->E, nullptr},
            {h->d_theta, sizeof(float) * (size_t)h->N, nullptr},
            {h->d_theta_hat, sizeof(float) * (size_t)h->N, nullptr},
            {h->d_omega_hat, sizeof(float) * (size_t)h->N, nullptr},
            {h->d_ring, h->ring_bytes, nullptr},
            {h->d_ctrl, sizeof(NetCtrl) * (size_t)h->n_nets, nullptr},
            {h->d_step, sizeof(int64_t), nullptr}};
    step = h->step;
    phi = h->phi;
    h->phi = nullptr;
    bool ok = hipStreamSynchronize(h->stream) == hipSuccess;
    for (auto& b : bufs) {
      if (!ok) break;
      ok = hipMalloc(&b.save, std::max<size_t>(b.bytes, 1)) == hipSuccess &&
           hipMemcpyAsync(b.save, b.dev, b.bytes, hipMemcpyDeviceToDevice, h->stream) == hipSuccess;
    }
    return ok;
  }
  bool restore() {
    bool ok = true;
    for (auto& b : bufs)
      ok = ok && hipMemcpyAsync(b.dev, b.save, b.bytes, hipMemcpyDeviceToDevice, h->stream) == hipSuccess;
    return ok;
  }
  // Restores the state, the step count and the hub statistics, and frees the copy.
  bool finish() {
    bool ok = restore() && hipStreamSynchronize(h->stream) == hipSuccess;
    for (auto& b : bufs)
      if (b.save) (void)hipFree(b.save);
    bufs.clear();
    h->step = step;
    h->phi = phi;
    return ok;
  }
};

// ms per step of `steps` steps (whole blocks of the multi-step scatter) from the snapshot, after one warm-up block
bool time_trial(LeviathanHandle* h, StateSnapshot& snap, float dt, int steps, hipEvent_t e0, hipEvent_t e1,
                double* per) {
  const int n = std::max(steps / block_steps(h), 1) * block_steps(h);
  bool ok = snap.restore() && enqueue_steps(h, dt, block_steps(h)) && hipEventRecord(e0, h->stream) == hipSuccess &&
            enqueue_steps(h, dt, n) && hipEventRecord(e1, h->stream) == hipSuccess &&
            hipEventSynchronize(e1) == hipSuccess;
  float ms = 0.f;
  if (ok) (void)hipEventElapsedTime(&ms, e0, e1);
  *per = ms / n;
  return ok;
}

bool resize_hbuf(LeviathanHandle* h, int steps) {
  if (steps <= h->hbuf_steps) return true;
  const size_t one = std::max<size_t>((size_t)h->E * (h->cfg.history == LEV_HISTORY_F32 ? 4 : 2), 1);
  void* bufs[2] = {nullptr, nullptr};
  const int nb = h->d_hbuf2 ? 2 : 1;  // the regrouped values of the staged schedule too
  for (int i = 0; i < nb; ++i) {
    const hipError_t e = hipMalloc(&bufs[i], one * (size_t)steps);
    if (e != hipSuccess) {
      set_err("hipMalloc(%zu bytes) for the values of %d steps failed: %s", one * (size_t)steps, steps,
              hipGetErrorString(e));
      if (bufs[0]) (void)hipFree(bufs[0]);
      return false;
    }
  }
  // the captured graphs hold the old arrays' addresses: drop them before the
  // arrays are freed (drop_graphs drains the stream first)
  drop_graphs(h);
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    set_err("hipStreamSynchronize failed");
    for (void* p : bufs)
      if (p) (void)hipFree(p);
    return false;
  }
  (void)hipFree(h->d_hbuf);
  h->d_hbuf = bufs[0];
  if (nb == 2) {
    (void)hipFree(h->d_hbuf2);
    h->d_hbuf2 = bufs[1];
  }
  h->vram += (size_t)nb * one * (size_t)(steps - h->hbuf_steps);
  h->hbuf_steps = steps;
  return true;
}

}  // namespace

int leviathan_autotune(LeviathanHandle* h, float dt, int steps, double* ms_out) {
  if (!check_handle(h)) return -1;
  if (steps <= 0) {
    set_err("steps must be positive");
    return -1;
  }
  if (h->drive_on) {
    set_err("autotune runs undriven steps; switch the drive off first");
    return -1;
  }
  StateSnapshot snap;
  bool ok = snap.take(h);
  const int active0 = h->active;
  double best = 1e300;
  int best_idx = active0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ok = ok && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
  for (int q = 0; q < kNumPlans; ++q) {
    if (ms_out) ms_out[q] = std::numeric_limits<double>::quiet_NaN();
    if (!ok || !h->plans[q].built) continue;
    h->active = q;
    double per = 0.0;
    ok = time_trial(h, snap, dt, steps, e0, e1, &per);
    if (ms_out) ms_out[q] = per;
    if (ok && per < best) {
      best = per;
      best_idx = q;
    }
  }
  // restore the state the call started from and keep the fastest plan
  ok = snap.finish() && ok;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (best_idx != active0) drop_graphs(h);
  h->active = ok ? best_idx : active0;
  if (!ok) {
    set_err("autotune: a HIP call failed");
    return -1;
  }
  return 0;
}

int leviathan_set_scatter_steps(LeviathanHandle* h, int steps) {
  if (!check_handle(h)) return -1;
  if (steps != 1 && steps != 2 && steps != lev::kMaxScatterSteps) {
    set_err("scatter steps must be 1, 2 or %d (got %d)", lev::kMaxScatterSteps, steps);
    return -1;
  }
  if (!h->d_hbuf) {
    set_err("no two-phase schedule was built for this handle");
    return -1;
  }
  if (steps > h->n_slots) {  // the pass reads slots t - d .. t + K - 1 - d of the S-slot ring
    set_err("%d steps per scatter pass exceed the ring's %d slots (max_delay + 2)", steps, h->n_slots);
    return -1;
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    set_err("hipStreamSynchronize failed");
    return -1;
  }
  if (!resize_hbuf(h, steps)) return -1;
  if (steps != h->scatter_steps) drop_graphs(h);  // captured graphs hold the old block structure
  h->scatter_steps = steps;
  return 0;
}

int leviathan_get_scatter_steps(LeviathanHandle* h) {
  if (!check_handle(h)) return -1;
  return h->scatter_steps;
}

int leviathan_set_row_steps(LeviathanHandle* h, int steps) {
  if (!check_handle(h)) return -1;
  if (steps != 1 && steps != 2) {
    set_err("row steps must be 1 or 2 (got %d)", steps);
    return -1;
  }
  if (steps == 2 && h->min_delay < 1) {
    set_err("paired row passes need every delay >= 1; this graph has delay-0 edges");
    return -1;
  }
  if (steps != h->row_steps) drop_graphs(h);  // captured graphs hold the old launch sequence
  h->row_steps = steps;
  return 0;
}

int leviathan_get_row_steps(LeviathanHandle* h) {
  if (!check_handle(h)) return -1;
  return h->row_steps;
}

int leviathan_get_layout(LeviathanHandle* h) {
  if (!check_handle(h)) return -1;
  return h->cfg.layout;
}

int leviathan_autotune_scatter_steps(LeviathanHandle* h, float dt, int steps, double* ms_out) {
  if (!check_handle(h)) return -1;
  if (steps <= 0) {
    set_err("steps must be positive");
    return -1;
  }
  if (h->drive_on) {
    set_err("autotune runs undriven steps; switch the drive off first");
    return -1;
  }
  const int cand[3] = {1, 2, lev::kMaxScatterSteps};
  for (int i = 0; i < 3; ++i)
    if (ms_out) ms_out[i] = std::numeric_limits<double>::quiet_NaN();
  if (h->plans[h->active].kind == LEV_SCHEDULE_DIRECT) return 0;  // the direct schedule has no scatter
  StateSnapshot snap;
  bool ok = snap.take(h);
  const int k0 = h->scatter_steps;
  double best = 1e300;
  int best_k = k0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ok = ok && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
  for (int i = 0; i < 3 && ok; ++i) {
    if (cand[i] > h->n_slots) continue;  // more steps than ring slots
    if (!resize_hbuf(h, cand[i])) {
      h->scatter_steps = k0;
      snap.finish();
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      return -1;
    }
    h->scatter_steps = cand[i];
    double per = 0.0;
    ok = time_trial(h, snap, dt, steps, e0, e1, &per);
    if (ms_out) ms_out[i] = per;
    if (ok && per < best) {
      best = per;
      best_k = cand[i];
    }
  }
  ok = snap.finish() && ok;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  h->scatter_steps = ok ? best_k : k0;
  if (h->scatter_steps != k0) drop_graphs(h);
  if (!ok) {
    set_err("autotune_scatter_steps: a HIP call failed");
    return -1;
  }
  return 0;
}

int leviathan_features_setup(LeviathanHandle* h, const int* nodes, int n_nodes, int max_rows) {
  if (!check_handle(h)) return -1;
  if (!nodes || n_nodes <= 0 || max_rows <= 0) {
    set_err("need feature nodes and max_rows > 0");
    return -1;
  }
  for (int f = 0; f < n_nodes; ++f)
    if (nodes[f] < 0 || nodes[f] >= h->N) {
      set_err("feature node %d out of range", nodes[f]);
      return -1;
    }
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    set_err("hipStreamSynchronize failed");
    return -1;
  }
  drop_graphs(h);  // captured graphs hold the old feature buffers
  void* old[] = {h->d_feat_nodes, h->d_X, h->d_feat_base, h->d_cs_partial};
  for (void* p : old)
    if (p) (void)hipFree(p);
  h->d_feat_nodes = nullptr;
  h->d_X = nullptr;
  h->d_feat_base = nullptr;
  h->d_cs_partial = nullptr;
  h->feat_n = n_nodes;
  h->feat_max_rows = max_rows;
  h->feat_rows = 0;
  const size_t F = 2 * (size_t)n_nodes;
  if (!dalloc(h, &h->d_feat_nodes, (size_t)n_nodes) || !dalloc(h, &h->d_X, (size_t)max_rows * F) ||
      !dalloc(h, &h->d_feat_base, 1) || !dalloc(h, &h->d_cs_partial, (size_t)lev::kColsumChunks * F)) {
    h->feat_n = 0;
    return -1;
  }
  if (hipMemcpy(h->d_feat_nodes, nodes, sizeof(int) * n_nodes, hipMemcpyHostToDevice) != hipSuccess) {
    set_err("copy feature nodes failed");
    return -1;
  }
  return 0;
}

int leviathan_step_n_features(LeviathanHandle* h, float dt, int n, const float* u, float* r_out) {
  if (!check_handle(h)) return -1;
  if (h->feat_n == 0) {
    set_err("call leviathan_features_setup first");
    return -1;
  }
  if (n < 0 || h->feat_rows + n > h->feat_max_rows) {
    set_err("feature rows would exceed max_rows (%lld + %d > %lld)", (long long)h->feat_rows, n,
            (long long)h->feat_max_rows);
    return -1;
  }
  if (u && !h->drive_on) {
    set_err("an input sequence needs a drive: call leviathan_set_drive first");
    return -1;
  }
  // rows are numbered from the step count at which capture of this batch starts
  const int64_t base = h->step - h->feat_rows;
  if (hipMemcpy(h->d_feat_base, &base, sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess) {
    set_err("copy feature base failed");
    return -1;
  }
  h->capturing = true;
  const int seg_max = std::min(lev::kRHistSteps, lev::kDriveSteps);
  std::vector<float> zeros;
  bool ok = true;
  for (int done = 0; done < n && ok;) {
    const int seg = std::min(n - done, seg_max);
    if (h->drive_on) {
      if (!u) zeros.assign(seg, 0.0f);
      ok = upload_drive_u(h, u ? u + done : zeros.data(), seg);
    }
    ok = ok && run_segment(h, dt, seg, r_out ? r_out + (size_t)done * h->n_nets : nullptr);
    done += seg;
    if (ok) h->feat_rows += seg;
  }
  h->capturing = false;
  return ok ? 0 : -1;
}

int64_t leviathan_features_rows(LeviathanHandle* h) { return h ? h->feat_rows : 0; }

namespace {
bool feat_range(LeviathanHandle* h, int64_t row0, int64_t nrows) {
  if (h->feat_n == 0) {
    set_err("no features: call leviathan_features_setup first");
    return false;
  }
  if (row0 < 0 || nrows <= 0 || row0 + nrows > h->feat_rows) {
    set_err("rows [%lld, %lld) outside the %lld captured rows", (long long)row0, (long long)(row0 + nrows),
            (long long)h->feat_rows);
    return false;
  }
  return true;
}
}  // namespace

int leviathan_features_get(LeviathanHandle* h, int64_t row0, int64_t nrows, float* out) {
  if (!check_handle(h) || !feat_range(h, row0, nrows)) return -1;
  const size_t F = 2 * (size_t)h->feat_n;
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess)
    e = hipMemcpy(out, h->d_X + (size_t)row0 * F, sizeof(float) * F * (size_t)nrows, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_err("features_get: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int leviathan_features_colsum(LeviathanHandle* h, int64_t row0, int64_t nrows, const double* weights, double* out) {
  if (!check_handle(h) || !feat_range(h, row0, nrows)) return -1;
  const int32_t F = 2 * h->feat_n;
  double *d_w = nullptr, *d_out = nullptr;
  hipError_t e = hipMalloc((void**)&d_out, sizeof(double) * F);
  if (e == hipSuccess && weights) e = hipMalloc((void**)&d_w, sizeof(double) * nrows);
  if (e == hipSuccess && weights) e = hipMemcpy(d_w, weights, sizeof(double) * nrows, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = lev::launch_colsum(h->d_X, F, row0, nrows, d_w, F, h->d_cs_partial, lev::kColsumChunks, d_out, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(out, d_out, sizeof(double) * F, hipMemcpyDeviceToHost);
  if (d_w) (void)hipFree(d_w);
  if (d_out) (void)hipFree(d_out);
  if (e != hipSuccess) {
    set_err("features_colsum: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int leviathan_features_colsum_noalloc(LeviathanHandle* h, int64_t row0, int64_t nrows, const double* weights,
                                      double* out) {
  if (!check_handle(h) || !feat_range(h, row0, nrows)) return -1;
  if (!out) {
    set_err("null output array");
    return -1;
  }
  const int32_t F = 2 * h->feat_n;
  if (F > kNoAllocFeatures || (weights && nrows > kNoAllocRows)) {
    set_err("features_colsum_noalloc: %d features and %lld weighted rows exceed the allocation-free maximum "
            "(%d features, %lld weighted rows); use leviathan_features_colsum",
            F, (long long)(weights ? nrows : 0), kNoAllocFeatures, (long long)kNoAllocRows);
    return -1;
  }
  double* base = scratch_ptr(h->device);
  if (!base) return -1;
  double* d_w = weights ? base : nullptr;
  double* d_out = base + kNoAllocRows;
  hipError_t e = hipSuccess;
  if (weights) e = hipMemcpyAsync(d_w, weights, sizeof(double) * nrows, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = lev::launch_colsum(h->d_X, F, row0, nrows, d_w, F, h->d_cs_partial, lev::kColsumChunks, d_out, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(double) * F, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    set_err("features_colsum_noalloc: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int leviathan_features_matvec(LeviathanHandle* h, int64_t row0, int64_t nrows, const double* mean, const double* w,
                              double* out) {
  if (!check_handle(h) || !feat_range(h, row0, nrows)) return -1;
  if (!mean || !w || !out) {
    set_err("null array");
    return -1;
  }
  const int32_t F = 2 * h->feat_n;
  double *d_mean = nullptr, *d_w = nullptr, *d_out = nullptr;
  hipError_t e = hipMalloc((void**)&d_mean, sizeof(double) * F);
  if (e == hipSuccess) e = hipMalloc((void**)&d_w, sizeof(double) * F);
  if (e == hipSuccess) e = hipMalloc((void**)&d_out, sizeof(double) * nrows);
  if (e == hipSuccess) e = hipMemcpy(d_mean, mean, sizeof(double) * F, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_w, w, sizeof(double) * F, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = lev::launch_matvec(h->d_X, F, row0, nrows, d_mean, d_w, F, d_out, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(out, d_out, sizeof(double) * nrows, hipMemcpyDeviceToHost);
  void* ptrs[] = {d_mean, d_w, d_out};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (e != hipSuccess) {
    set_err("features_matvec: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int leviathan_features_gram(LeviathanHandle* h, int64_t row0, int64_t nrows, const double* mean, int terms,
                            float* gram_out, double* ms_out) {
  if (!check_handle(h) || !feat_range(h, row0, nrows)) return -1;
  if (!mean || !gram_out || (terms != 1 && terms != 3)) {
    set_err("need mean and output arrays and terms 1 or 3");
    return -1;
  }
  const int32_t F = 2 * h->feat_n;
  float* d_g = nullptr;
  if (!gram_on_device(h, row0, nrows, mean, terms, &d_g, ms_out)) return -1;
  const hipError_t e = hipMemcpy(gram_out, d_g, sizeof(float) * (size_t)F * F, hipMemcpyDeviceToHost);
  (void)hipFree(d_g);
  if (e != hipSuccess) {
    set_err("features_gram: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

int leviathan_features_factor(LeviathanHandle* h, int64_t row0, int64_t nrows, const double* mean, int terms,
                              double alpha, double* shift_out, double* gram_ms_out) {
  if (!check_handle(h) || !feat_range(h, row0, nrows)) return -1;
  if (!mean || (terms != 0 && terms != 1 && terms != 3) || !(alpha >= 0.0)) {
    set_err("need the mean, terms 0, 1 or 3 and alpha >= 0");
    return -1;
  }
  const int32_t F = 2 * h->feat_n;
  if (!h->blas) {
    if (rocblas_create_handle(&h->blas) != rocblas_status_success) {
      h->blas = nullptr;
      set_err("features_factor: rocblas_create_handle failed");
      return -1;
    }
    (void)rocblas_set_stream(h->blas, h->stream);
  }
  if (h->chol_n != F) {
    void* rp[] = {h->d_chol, h->d_vec, h->d_dmean, h->d_info};
    for (void* p : rp)
      if (p) (void)hipFree(p);
    h->d_chol = nullptr;
    h->d_vec = h->d_dmean = nullptr;
    h->d_info = nullptr;
    h->chol_n = 0;
    hipError_t e = hipMalloc((void**)&h->d_chol, sizeof(double) * (size_t)F * F);
    if (e == hipSuccess) e = hipMalloc((void**)&h->d_vec, sizeof(double) * 2 * F);
    if (e == hipSuccess) e = hipMalloc((void**)&h->d_dmean, sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&h->d_info, sizeof(rocblas_int));
    if (e != hipSuccess) {
      set_err("features_factor: %s", hipGetErrorString(e));
      return -1;
    }
    h->chol_n = F;
  }
  // the Gram in FP64: widened from the matrix-core Gram (terms 1, 3) or formed in FP64 (terms 0)
  double* d_c = nullptr;
  hipError_t e = hipMalloc((void**)&d_c, sizeof(double) * (size_t)F * F);
  if (e != hipSuccess) {
    set_err("features_factor: %s", hipGetErrorString(e));
    return -1;
  }
  auto fail = [&](const char* what) {
    (void)hipFree(d_c);
    set_err("features_factor: %s", what);
    return -1;
  };
  if (terms != 0) {
    float* d_g = nullptr;
    if (!gram_on_device(h, row0, nrows, mean, terms, &d_g, gram_ms_out)) {
      (void)hipFree(d_c);
      return -1;
    }
    e = lev::launch_widen(d_g, (int64_t)F * F, d_c, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d_g);
    if (e != hipSuccess) return fail(hipGetErrorString(e));
  } else {
    double* d_mean = nullptr;
    double* d_xc = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    e = hipMalloc((void**)&d_mean, sizeof(double) * F);
    if (e == hipSuccess) e = hipMalloc((void**)&d_xc, sizeof(double) * (size_t)nrows * F);
    if (e == hipSuccess) e = hipMemcpy(d_mean, mean, sizeof(double) * F, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = lev::launch_center_fp64(h->d_X, F, row0, nrows, d_mean, F, d_xc, h->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
    bool blas_ok = e == hipSuccess;
    if (blas_ok) {
      // C = Xc^T Xc: Xc is [nrows][F] row-major, i.e. column-major F x nrows
      const double one = 1.0, zero = 0.0;
      blas_ok = rocblas_dsyrk(h->blas, rocblas_fill_lower, rocblas_operation_none, F, (rocblas_int)nrows, &one, d_xc,
                              F, &zero, d_c, F) == rocblas_status_success;
    }
    if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && gram_ms_out) {
      float ms = 0.f;
      e = hipEventElapsedTime(&ms, e0, e1);
      *gram_ms_out = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (d_mean) (void)hipFree(d_mean);
    if (d_xc) (void)hipFree(d_xc);
    if (e != hipSuccess) return fail(hipGetErrorString(e));
    if (!blas_ok) return fail("rocblas_dsyrk failed");
  }
  double scale = 0.0;
  e = lev::launch_diag_mean(d_c, F, h->d_dmean, h->stream);
  if (e == hipSuccess) e = hipMemcpy(&scale, h->d_dmean, sizeof(double), hipMemcpyDeviceToHost);
  if (!(scale > 0.0)) scale = 1.0;
  double shift = 0.0;
  bool done = false;
  for (int attempt = 0; e == hipSuccess && attempt < 16 && !done; ++attempt) {
    e = lev::launch_shifted(d_c, F, alpha + shift, h->d_chol, h->stream);
    if (e != hipSuccess) break;
    if (rocsolver_dpotrf(h->blas, rocblas_fill_lower, F, h->d_chol, F, h->d_info) != rocblas_status_success)
      return fail("rocsolver_dpotrf failed");
    rocblas_int info = 0;
    e = hipMemcpy(&info, h->d_info, sizeof(rocblas_int), hipMemcpyDeviceToHost);
    if (e == hipSuccess && info == 0) {
      // For the approximate (matrix-core) Grams, the inverse from the factor: each
      // preconditioner application is then one bandwidth-bound DSYMV rather than two
      // latency-bound triangular solves.  The FP64 Gram serves ill-conditioned features,
      // where the explicit inverse would lose digits: it keeps the triangular solves.
      h->chol_inverse = terms != 0;
      if (h->chol_inverse) {
        if (rocsolver_dpotri(h->blas, rocblas_fill_lower, F, h->d_chol, F, h->d_info) != rocblas_status_success)
          return fail("rocsolver_dpotri failed");
        e = hipMemcpy(&info, h->d_info, sizeof(rocblas_int), hipMemcpyDeviceToHost);
        if (e == hipSuccess && info != 0) {
          // the inverse of a nearly singular factor did not form: factorise again and keep the
          // triangular solves
          h->chol_inverse = false;
          e = lev::launch_shifted(d_c, F, alpha + shift, h->d_chol, h->stream);
          if (e == hipSuccess &&
              rocsolver_dpotrf(h->blas, rocblas_fill_lower, F, h->d_chol, F, h->d_info) != rocblas_status_success)
            return fail("rocsolver_dpotrf failed");
          if (e == hipSuccess) e = hipMemcpy(&info, h->d_info, sizeof(rocblas_int), hipMemcpyDeviceToHost);
          if (e == hipSuccess && info != 0) return fail("the factorisation did not repeat");
        }
      }
      done = true;
    } else {
      shift = shift == 0.0 ? 1e-7 * scale : shift * 10.0;  // not positive definite: a larger shift
    }
  }
  if (e != hipSuccess) return fail(hipGetErrorString(e));
  (void)hipFree(d_c);
  if (!done) {
    set_err("features_factor: no diagonal shift up to %g made the Gram factorisable", shift);
    return -1;
  }
  if (shift_out) *shift_out = shift;
  return 0;
}

int leviathan_features_precond(LeviathanHandle* h, const double* r, double* z) {
  if (!check_handle(h)) return -1;
  if (!h->d_chol || h->chol_n != 2 * h->feat_n) {
    set_err("features_precond: no factor (call leviathan_features_factor first)");
    return -1;
  }
  if (!r || !z) {
    set_err("features_precond: null vector");
    return -1;
  }
  const int32_t F = h->chol_n;
  hipError_t e = hipMemcpyAsync(h->d_vec, r, sizeof(double) * F, hipMemcpyHostToDevice, h->stream);
  const double one = 1.0, zero = 0.0;
  if (h->chol_inverse) {
    if (e == hipSuccess && rocblas_dsymv(h->blas, rocblas_fill_lower, F, &one, h->d_chol, F, h->d_vec, 1, &zero,
                                         h->d_vec + F, 1) != rocblas_status_success) {
      set_err("features_precond: rocblas_dsymv failed");
      return -1;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(z, h->d_vec + F, sizeof(double) * F, hipMemcpyDeviceToHost, h->stream);
  } else {
    if (e == hipSuccess &&
        rocsolver_dpotrs(h->blas, rocblas_fill_lower, F, 1, h->d_chol, F, h->d_vec, F) != rocblas_status_success) {
      set_err("features_precond: rocsolver_dpotrs failed");
      return -1;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(z, h->d_vec, sizeof(double) * F, hipMemcpyDeviceToHost, h->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    set_err("features_precond: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

}  // extern "C"
