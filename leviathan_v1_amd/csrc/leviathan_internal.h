// leviathan_internal.h -- device-side data structures shared by the HIP
// kernels (leviathan_kernels.hip) and the host C-ABI (leviathan_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "phi_internal.h"

namespace lev {

constexpr int kBlock = 256;          // lanes per step workgroup (4 waves of 64)
constexpr int kFinalizeThreads = 1024;
constexpr int kRHistSteps = 4096;    // steps of r kept on the device ring
constexpr int kDriveSteps = 4096;    // steps of drive input u_t kept on the device ring
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, workgroups dealt round-robin

// Row block flag: a single row whose in-degree exceeds the block capacity.
constexpr int kBlockLongRow = 1;

// Row-pass work item: {first row, first edge, end edge, net << 10 | rows << 1 | long}.
__host__ __device__ inline int4 pack_item(int32_t r0, int32_t e0, int32_t e1, int32_t nrows, int32_t net,
                                          bool is_long) {
  return make_int4(r0, e0, e1, (net << 10) | (nrows << 1) | (is_long ? kBlockLongRow : 0));
}
constexpr int kFinishBlocks = 8;     // row blocks per workgroup of the paired pass's node pass
constexpr int kSerialRow = 128;      // stream blocks: longer rows are filled and summed by a wavefront
// Estimated row-pass cost of a block (ns of workgroup lifetime), for balancing the XCDs.
constexpr double kCostBlock = 9860.0, kCostEdge = 7.42, kCostRow = 11.9, kCostHeavyRow = 300.0;
constexpr double kCostLong = 20900.0, kCostLongEdge = 5.59;

// Per-network controller state, read by the step kernel (g) and updated by
// the finalize kernel (r, g) in FP64.
struct NetCtrl {
  double g;        // gain used by the next step
  double r;        // r of the last step
  double r0, beta, mu, g_rest, g_min, g_max;
  int32_t controller;
  int32_t n_nodes;
};

template <class H>
struct RingRef {
  H* ring;
  int64_t pitch;  // node-major: elements per node; slot-major: elements per slot
};

struct StepArgs {
  const int32_t* row_ptr;
  const int32_t* col;
  const uint8_t* delay;
  float* w;
  float* theta;
  float* theta_hat;
  const float* omega;
  float* omega_hat;
  void* ring;
  int64_t ring_pitch;
  int32_t n_slots;        // S = max_delay + 2
  int32_t num_items;        // row blocks (long ones first, then stream ones), one workgroup each
  // items [xcd_split[x], xcd_split[x+1]) run on XCD x: workgroup 8 * j + x takes item xcd_split[x] + j
  int32_t xcd_split[kNumXcd + 1];
  const int4* items;        // pack_item layout
  double2* partials;        // per item: (sum cos, sum sin) of the new phases
  const NetCtrl* ctrl;
  const int64_t* step;    // global step counter (device)
  const float* drive_b;   // NULL: no input drive
  const float* drive_u;   // [kDriveSteps] ring, u_t at t % kDriveSteps
  // binned schedule: delayed phases written by the scatter kernel; staged: the regrouped values of
  // the step (written here only by a paired pass gathering a long row's early edges)
  void* hbuf;
  const uint32_t* hpos;   // binned schedule: [E] position of each CSR edge's value in hbuf
  int32_t hstride;        // binned schedule: elements per hbuf slot (the K interleaved steps of a scatter pass)
  const uint16_t* lpos;   // staged schedule: [E] position of each CSR edge's value in its block's range
  // paired row pass (two steps per pass): the staged values of the second step, and per node the
  // second step's coupling sum and prediction error, left for its integration
  void* hbuf_next;
  float2* pair;
  // staged schedule, the second pair of a multi-step block (early_k = its first step's index k > 0):
  // each row block's early edges, early[early_ptr[b] .. early_ptr[b + 1]) = {source, position in the
  // block's range << 2 | delay}; the pass gathers step k's value of those with delay < k and step
  // k + 1's of all of them from the ring
  const int32_t* early_ptr;
  const int2* early;
  int32_t early_k;
  float dt, eta, lam, gamma, alpha, kappa_w, w_min, w_max;
};

struct FinalizeArgs {
  const double2* partials;
  NetCtrl* ctrl;
  int64_t* step;
  float* r_hist;          // [kRHistSteps][n_nets]
  const int4* ranges;     // per network: [begin, end) of its long and of its stream items' partials
  int32_t n_nets;
  int32_t num_blocks;     // partials of the active plan
  float dt;
  const float* theta;     // phases after the step (read by the hub accumulation)
  PhiDev phi;             // phi.st null: no hub statistics attached
  double2* chunks;        // [ceil(num_blocks / kChunkPartials)] chunk sums (one network), or null
  unsigned int* fin_done; // several networks: workgroups of the finalize grid that have finished (0 between steps)
};

// One network with many row blocks: a grid of workgroups first sums chunks of
// kChunkPartials partials (one round of loads each), and the finalize kernel
// sums the chunk sums.  A single workgroup reading 10^5 partials is bound by
// load latency (~30 us on the 10^7-node graph).
constexpr int kChunkThreads = 256;
constexpr int kChunkPartials = kChunkThreads * 8;
constexpr int kChunkMinBlocks = 16384;

// Two-phase (binned) step: see leviathan_binned.hip.
constexpr int kScatterEpt = 16;     // phase 1: out-edges per lane in a source block

// Regroup pass of the staged schedule (leviathan_binned.hip).
constexpr int kRegroupChunk = 4096;  // bucket slots per regroup workgroup
constexpr int kRegroupRuns = 1024;   // runs per chunk
constexpr int kRegroupChunksPerWg = 2;  // chunks one regroup workgroup loads at once
struct RegroupArgs {
  const int4* chunks;        // {first slot, count, first run, run count}
  const uint16_t* rank;      // [E] bucket slot -> index in its chunk's staged order
  const int2* runs;          // {first index in the chunk's staged order, staged slot}
  const void* src;           // bucketed values, [E][steps] (a vector of the pass's steps per slot)
  void* dst;                 // staged values, [steps][stride] (pack_pairs: [steps / 2][stride] pairs)
  int32_t num_chunks;
  int64_t stride;            // slots between consecutive steps' (or pairs') staged arrays (steps > 1)
  int32_t steps;             // steps whose values one pass regroups (1, 2 or kMaxScatterSteps)
  int32_t pack_pairs;        // uint16 values, paired row passes: steps 2p, 2p + 1 as one 4-byte pair
};
hipError_t launch_regroup(int history, const RegroupArgs& a, hipStream_t s);

// The scatter's window form (slot-major rings): a short source block's <= kBlock
// sources start at most 7 columns past a multiple of 8, so kScatterWinCols
// columns of every slot row hold them; the window of all S slots must fit
// kScatterWinBytes of LDS.
constexpr int kScatterWinCols = kBlock + 8;
constexpr int kScatterWinBytes = 32768;
__host__ __device__ constexpr int scatter_window_slots(size_t elem_bytes) {
  return (int)(kScatterWinBytes / (kScatterWinCols * elem_bytes));
}
// The window form is used for slot-major rings (layout 1) that fit; its short source
// blocks store their out-edges in value-slot order with one code per edge
// (source in block << 8 | delay, run index << 16) in place of the rank and run.
__host__ __device__ constexpr bool scatter_uses_window(int layout, int n_slots, size_t elem_bytes) {
  return layout == 1 && n_slots <= scatter_window_slots(elem_bytes);
}
constexpr int kScatterRuns = 1024;  // runs of a short source block whose slot bases are staged in LDS
static_assert(kScatterRuns % kBlock == 0, "run bases are staged kBlock at a time");
struct ScatterArgs {
  const int32_t* csc_ptr;   // [N+1] out-edge offsets per source
  const uint8_t* delay;     // [E] delays in CSC order
  const uint32_t* pos;      // [E] long sources: value slots in CSC order
  // [E] short source blocks: CSC edge -> its index in the block's slot order (low 16 bits) and each
  // slot-order index's run within its block (high 16 bits)
  const uint32_t* rank_run;
  const int32_t* run_ptr;   // [num_blocks + 1] each block's runs of consecutive slots
  const int32_t* run_base;  // slot-order index k of run r goes to slot run_base[r] + k
  const uint16_t* dsrc;     // [E] slot-major ring: each edge's source index in its block << 8 | delay
  const void* ring;
  int64_t ring_pitch;
  int64_t ring_bytes;       // size of the ring (a node-major ring below 2^32 bytes is read through 32-bit offsets)
  int32_t n_slots;
  int32_t num_blocks;
  int32_t xcd_remap;
  const int4* blocks;       // pack_item layout: {first source, first out-edge, end out-edge, sources << 1 | long}
  void* hbuf;               // [E][steps] delayed phases, bucketed: each slot holds the values of the
                            // pass's steps side by side (one vector per edge)
  int32_t steps;            // 1, 2 or kMaxScatterSteps steps gathered in one pass
  const int64_t* step;
};

// One scatter pass can gather the values of up to kMaxScatterSteps consecutive
// steps t..t+K-1: step t+k reads theta(t+k-d), which exists before step t's
// row pass whenever d >= k.  The remaining edges (d < k) are gathered by
// early_kernel just before step t+k's row pass, from a list of the edges with
// delay < kMaxScatterSteps - 1 sorted by (delay, source).
constexpr int kMaxScatterSteps = 4;
struct EarlyArgs {
  const int32_t* src;       // [count] source node
  const uint32_t* pos;      // [count] value slot in that array
  const uint8_t* delay;     // [count] delay, ascending
  int32_t count;            // the records with delay < k
  const void* ring;
  int64_t ring_pitch;
  int32_t n_slots;
  void* hbuf;               // the values the step's row pass reads (bucketed, or regrouped when staged)
  int32_t hstride;          // elements per slot of that array (binned: the pass's interleaved steps)
  const int64_t* step;
  int32_t step_off;         // the values are those of step *step + step_off (1: gathered a step ahead,
                            // for the second step of a paired row pass)
};

// The delayed values of a scatter pass's K steps for one edge, side by side
// (4 or 8 bytes for uint16 history): one LDS access and one store per edge
// and pass instead of one per step.
template <class H, int K>
struct alignas(sizeof(H) * K) HVec {
  H x[K];
};

// Launchers (leviathan_kernels.hip).  Each returns the launch status.
// gather: 0 = delayed phases read from the history ring (direct schedule),
//         1 = read from the scatter kernel's bucketed array (binned schedule)
hipError_t launch_step(int history, int layout, int ept, int gather, const StepArgs& a, hipStream_t s);
// The paired row pass: steps t and t+1 of every row block but step t+1's
// integration, which launch_pair_finish runs after step t's finalize.  Needs a
// graph without delay-0 edges, a.pair, and (staged) a.hbuf_next.
hipError_t launch_pair(int history, int layout, int ept, int gather, const StepArgs& a, hipStream_t s);
hipError_t launch_pair_finish(int history, int layout, const StepArgs& a, hipStream_t s);
hipError_t launch_finalize(const FinalizeArgs& a, hipStream_t s);
hipError_t launch_fill_ring(int history, int layout, void* ring, int64_t pitch, int32_t n_slots,
                            const float* theta, int32_t n, hipStream_t s);
hipError_t launch_write_slot(int history, int layout, void* ring, int64_t pitch, int32_t slot, const float* in,
                             int32_t n, hipStream_t s);
hipError_t launch_read_slot(int history, int layout, const void* ring, int64_t pitch, int32_t slot,
                            float* out, int32_t n, hipStream_t s);
// (leviathan_binned.hip)
hipError_t launch_scatter(int history, int layout, const ScatterArgs& a, hipStream_t s);
hipError_t launch_early(int history, int layout, const EarlyArgs& a, hipStream_t s);
// (reservoir.hip)
hipError_t launch_capture(const float* theta, const int32_t* nodes, int32_t n_nodes, float* X, int64_t max_rows,
                          const int64_t* step, const int64_t* base, hipStream_t s);
hipError_t launch_colsum(const float* X, int64_t ld, int64_t row0, int64_t nrows, const double* w, int32_t F,
                         double* partial, int32_t chunks, double* out, hipStream_t s);
hipError_t launch_matvec(const float* X, int64_t ld, int64_t row0, int64_t nrows, const double* mean,
                         const double* w, int32_t F, double* out, hipStream_t s);
hipError_t launch_center_bf16(const float* X, int64_t ld, int64_t row0, int64_t nrows, const double* mean,
                              int32_t F, int32_t Fp, uint16_t* XcT, int64_t Tp, uint16_t* XcT_lo, hipStream_t s);
hipError_t launch_gram(const uint16_t* XcT, int64_t Tp, int32_t Fp, int32_t F, float* G, hipStream_t s);
hipError_t launch_gram_split(const uint16_t* Hi, const uint16_t* Lo, int64_t Tp, int32_t Fp, int32_t F, float* G,
                             hipStream_t s);
hipError_t launch_widen(const float* G, int64_t n, double* C, hipStream_t s);
hipError_t launch_shifted(const double* C, int32_t F, double diag_add, double* A, hipStream_t s);
hipError_t launch_diag_mean(const double* C, int32_t F, double* diag_mean, hipStream_t s);
hipError_t launch_center_fp64(const float* X, int64_t ld, int64_t row0, int64_t nrows, const double* mean,
                              int32_t F, double* Xc, hipStream_t s);
constexpr int kColsumChunks = 128;

}  // namespace lev
