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
  std::vector<int4> sources;        // phase-1 work items {source_begin, source_end, 0, flags}
  std::vector<int32_t> csc_ptr;     // [N+1] out-edge offsets
  std::vector<uint8_t> delay_csc;   // [E] delays in CSC order
  std::vector<uint32_t> pos;        // [E] CSC edge -> slot in the bucketed array
  std::vector<uint32_t> hpos;       // [E] CSR edge -> slot in the bucketed array
  int32_t n_buckets = 0;
};

// Row blocks of the direct schedule: runs of <= 256 rows holding <= cap edges,
// or a single row with more than cap edges (listed first within its network,
// heaviest first).
bool build_direct_plan(const int* row_ptr, int n_nets, const int* net_offsets, int cap, HostPlan& out,
                       std::string& err);

// Buckets: runs of consecutive rows with <= bucket_edges incoming edges (a
// heavier row is a bucket by itself); inside a bucket the edges are ordered
// by (source, CSR index).  Also builds the source-ordered (CSC) scatter plan.
bool build_binned_plan(int N, const int* row_ptr, const int* col_idx, const uint8_t* delays, int n_nets,
                       const int* net_offsets, int bucket_edges, HostBinned& out, std::string& err);

}  // namespace lev
