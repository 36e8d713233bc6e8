// plan.h -- host-side work decomposition of the step (built once per handle).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace lev {

// Work items of one launch, in launch order, and each network's range of them
// (a network's items are contiguous; its partial sums are reduced over that range).
struct HostPlan {
  std::vector<int4> items;
  std::vector<int2> ranges;
};

// The two-phase (binned) schedule.
struct HostBinned {
  HostPlan buckets;                 // {row_begin, row_end, net, flags}
  std::vector<int2> sb_index;       // per bucket: {first bound, number of row sub-blocks}
  std::vector<int32_t> sb_bounds;   // row boundaries of the sub-blocks
  std::vector<int4> sources;        // phase-1 work items {source_begin, source_end, 0, flags}
  std::vector<int32_t> csc_ptr;     // [N+1] out-edge offsets
  std::vector<uint8_t> delay_csc;   // [E] delays in CSC order
  std::vector<uint32_t> pos;        // [E] bucketed position of each CSC edge
  std::vector<uint16_t> perm;       // [E] CSR edge -> position inside its bucket
  int32_t max_bucket_edges = 0;
};

// Row blocks of the direct schedule: runs of <= 256 rows holding <= cap edges,
// or a single row with more than cap edges (listed first within its network,
// heaviest first).
bool build_direct_plan(const int* row_ptr, int n_nets, const int* net_offsets, int cap, HostPlan& out,
                       std::string& err);

// Buckets of <= bucket_edges incoming edges (rows heavier than one sub-block
// become single-row buckets), their sub-blocks, the per-edge permutation and
// the source-ordered (CSC) scatter plan.
bool build_binned_plan(int N, const int* row_ptr, const int* col_idx, const uint8_t* delays, int n_nets,
                       const int* net_offsets, int bucket_edges, HostBinned& out, std::string& err);

}  // namespace lev
