// plan.h -- host-side work decomposition of the step (built once per handle).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace lev {

// Work items of the row pass (row blocks), in launch order: every network's
// long rows (heaviest first within a network, so the longest start first),
// then every network's stream blocks in CSR order.  An item is
// {r0, e0, e1, net << 10 | nrows << 1 | long} (pack_item), so a workgroup
// knows its edge range without reading row_ptr first; its partial sums sit at
// its index.  ranges[net] = {long_begin, long_end, stream_begin, stream_end}.
//
// xcd_split[0..8] cuts the items into 8 contiguous runs of about equal
// estimated cost, one per XCD (workgroups are dealt to the XCDs round-robin).
struct HostPlan {
  std::vector<int4> items;
  std::vector<int4> ranges;
  std::vector<int32_t> xcd_split;
};

// Phase 1 of the two-phase schedules: the graph by source.
struct HostScatter {
  std::vector<int4> sources;        // work items, pack_item layout {s0, c0, c1, nsrc << 1 | long}
  std::vector<int32_t> csc_ptr;     // [N+1] out-edge offsets
  std::vector<uint8_t> delay_csc;   // [E] delays in CSC order
  std::vector<uint32_t> csc_edge;   // [E] CSC index -> CSR edge
  std::vector<uint16_t> dsrc;       // [E] after order_scatter_by_delay: source index in the block << 8 | delay
};

// Row blocks of the direct schedule: runs of <= 256 rows holding <= cap edges,
// or a single row with more than cap edges.
bool build_direct_plan(const int* row_ptr, int n_nets, const int* net_offsets, int cap, HostPlan& out,
                       std::string& err);

// CSC of the graph and the phase-1 work items: runs of <= 256 sources with
// <= kBlock * kScatterEpt out-edges, or one heavier source.
bool build_scatter_plan(int N, const int* row_ptr, const int* col_idx, const uint8_t* delays, HostScatter& out,
                        std::string& err);

// Binned schedule: buckets are runs of consecutive rows with <= bucket_edges
// incoming edges (a heavier row is a bucket by itself); inside a bucket the
// values are ordered by (source, CSR index).  hpos[e] = slot of CSR edge e.
bool bucket_positions(int N, const int* row_ptr, const int* col_idx, int bucket_edges, std::vector<uint32_t>& hpos,
                      int32_t& n_buckets, std::string& err);

// Staged schedule: each row block of `plan` owns its CSR edge range; a
// stream block's values are ordered by (source, CSR index) inside it, a long
// row's keep CSR order.  hpos[e] = slot of CSR edge e, lpos[e] = hpos[e] - the
// block's first edge (stream blocks only).
bool staged_positions(int N, const int* row_ptr, const int* col_idx, const HostPlan& plan,
                      std::vector<uint32_t>& hpos, std::vector<uint16_t>& lpos, std::string& err);

// Staged schedule, phase-1 target: buckets are runs of consecutive row blocks
// of `plan` (in row order) with <= bucket_edges edges, each occupying its CSR
// edge range, values ordered by (source, CSR index) inside the bucket.
// hpos[e] = bucket slot of CSR edge e; bucket_begin lists each bucket's first
// slot (and E at the end).
bool staged_bucket_positions(int N, const int* row_ptr, const int* col_idx, const HostPlan& plan, int bucket_edges,
                             std::vector<uint32_t>& hpos, std::vector<int32_t>& bucket_begin, std::string& err);

// The regroup pass moves every value from its bucket slot to its staged slot.
// A bucket is cut into chunks of <= kRegroupChunk consecutive slots holding
// <= kRegroupRuns runs; a chunk's values for one row block are one run of
// consecutive staged slots (the staged region keeps source order), so the
// chunk is permuted in LDS into staged-slot order (rank) and leaves run by
// run.  chunks[k] = {first slot, count, first run, run count}; runs[r] =
// {first index in the chunk's staged order, its staged slot}.
struct HostRegroup {
  std::vector<int4> chunks;
  std::vector<uint16_t> rank;  // [E] per bucket slot
  std::vector<int2> runs;
};
bool build_regroup(const std::vector<uint32_t>& bucket_hpos, const std::vector<uint32_t>& staged_hpos,
                   const std::vector<int32_t>& bucket_begin, HostRegroup& out, std::string& err);

// pos[c] = hpos[CSR edge of CSC index c], checked to be a permutation of [0, E).
bool csc_positions(const HostScatter& sc, const std::vector<uint32_t>& hpos, std::vector<uint32_t>& pos,
                   std::string& err);

// Write order of the scatter kernel: inside each short source block the
// values are written in ascending slot order, so the block's values for one
// bucket or region (a contiguous run of slots, as slots are ordered by
// source there) leave as one contiguous run.  On return pos holds, per short
// block, the slots in ascending order, and rank[c] the index of CSC edge c in
// that order; long sources keep CSC order (rank unused).
void sort_scatter_writes(const HostScatter& sc, std::vector<uint32_t>& pos, std::vector<uint16_t>& rank);

// Runs of consecutive slots in each short source block's slot order (after
// sort_scatter_writes): slot-order index k belongs to run run_idx[c0 + k] of
// its block and goes to slot run_base[run_ptr[block] + run_idx] + k, so the
// scatter kernel reads 2 bytes per value instead of a 32-bit slot.
struct HostScatterRuns {
  std::vector<uint16_t> run_idx;  // [E]
  std::vector<int32_t> run_ptr;   // [source items + 1]
  std::vector<int32_t> run_base;  // [runs]
};
void encode_scatter_runs(const HostScatter& sc, const std::vector<uint32_t>& pos, HostScatterRuns& out);

// For a slot-major history ring: inside each short source block, order the
// out-edges by (delay, source, destination) -- delay_csc and csc_edge are
// permuted in place -- so that the lanes of one scatter load read one slot row
// for consecutive sources, and record each edge's source index in its block
// together with its delay.
// Must run before the positions are derived from csc_edge.
void order_scatter_by_delay(HostScatter& sc);

// Window form of the scatter (slot-major ring, after order_scatter_by_delay,
// sort_scatter_writes and encode_scatter_runs): inside each short source
// block the out-edges are stored in the order of their value slots, one code
// per edge: its source in the block << 8 | delay (low half) and its run of
// consecutive slots (high half).  Long sources keep CSC order (code unused).
void window_scatter_codes(const HostScatter& sc, const std::vector<uint16_t>& rank, const HostScatterRuns& runs,
                          std::vector<uint32_t>& code);

// Multi-step scatter: the CSR edges with delay < below, sorted by (delay,
// source, CSR index).  src/delay/csr hold each record's source, delay and CSR
// edge; cnt[d] (d = 0..below) = the number of records with delay < d.
void early_edges(int N, const int* row_ptr, const int* col_idx, const uint8_t* delays, int below,
                 std::vector<int32_t>& src, std::vector<uint8_t>& delay, std::vector<uint32_t>& csr,
                 std::vector<int32_t>& cnt);

// Paired row passes of the staged schedule: each row block's early edges
// (the records of early_edges), so that the second pair of a multi-step block
// gathers their values itself.  ptr[q] .. ptr[q + 1] are item q's records
// {source, (staged slot - the item's first edge) << 2 | delay}, in the order
// of the early list; delays must be below 4.
bool block_early_edges(const HostPlan& plan, const std::vector<int32_t>& src, const std::vector<uint8_t>& delay,
                       const std::vector<uint32_t>& csr, const std::vector<uint32_t>& staged_hpos,
                       std::vector<int32_t>& ptr, std::vector<int2>& rec, std::string& err);

}  // namespace lev
